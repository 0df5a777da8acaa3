"""Shared camera/light/shading setups for the tests."""
import math

import csgray as C
import oracle as O
from csgray import scenes as S

# The reference's defaults: Camera() at (0,0,5) looking down -z with a 90
# degree fov (Camera.h:18), DirectionalLight polar -60 / azimuth -45 degrees
# (DirectionalLight.h:10-11), Phong 0.2/0.8/0.7/30 (RaycastingKernels.cu:71-74).
DEFAULT_CAM6 = (0.0, 0.0, 5.0, 0.0, 0.0, 90.0 * 3.14159 / 180.0)
DEFAULT_PHONG = (0.2, 0.8, 0.7, 30.0)


def default_light3():
    p, a = -60.0 * 3.14159 / 180.0, -45.0 * 3.14159 / 180.0
    return (math.sin(p) * math.cos(a), math.cos(p), math.sin(p) * math.sin(a))


def view_camera():
    """Eye (3, 2.5, 4) looking at the origin, 60 degree vertical fov."""
    return C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)


def setups():
    """(name, csg camera, csg light, csg phong, oracle cam6, light3, phong8) pairs used across tests."""
    dl = default_light3()
    dcam = C.camera(DEFAULT_CAM6[:3], DEFAULT_CAM6[3], DEFAULT_CAM6[4], DEFAULT_CAM6[5])
    vcam = view_camera()
    return [
        ("default", dcam, C.light(dl), C.phong(*DEFAULT_PHONG), DEFAULT_CAM6, dl, O.phong8(*DEFAULT_PHONG)),
        ("view", vcam, C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG), C.cam_tuple(vcam), S.LIGHT_TOWARD,
         O.phong8(*S.PHONG)),
    ]
