"""Camera.MaxLightCount (camera.go:1474-1497) through the GPU path against
the oracle, bit for bit: the camera or a focus point as the reference, cuts
below and above the light count, both execution forms, and a light list with
several ambient and directional lights (whose comparator is not a strict
order, so the result rests on the stable sort's exact algorithm on both
sides)."""
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import AmbientLight, DirectionalLight
from tests.parity import compare_frames

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("focus", [None, (30.0, -10.0, -150.0)], ids=["camera", "focus_point"])
@pytest.mark.parametrize("count", [1, 3, 6, 30])
def test_max_light_count(gpu_ctx, oracle, count, focus, mode):
    wl = W.mixed_scene(count=12)
    wl.camera.max_light_count = count
    wl.camera.max_light_count_focus = focus
    print("max light count", count, "focus", focus, "mode", mode, flush=True)
    compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))


def test_max_light_count_with_tied_lights(gpu_ctx, oracle):
    wl = W.mixed_scene(count=12)
    extra = [AmbientLight("amb2", 0.2, 0.3, 0.4, 0.5), DirectionalLight("sun2", 1, 0.5, 0.5, 0.7),
             AmbientLight("amb3", 0.1, 0.1, 0.1, 0.3)]
    extra[1].set_local_rotation(W.axis_angle_matrix((0, 1, 0), 0.8))
    lights = wl.lights[:3] + [extra[0]] + wl.lights[3:6] + extra[1:] + wl.lights[6:]
    wl = W.Workload("tied_lights", wl.camera, wl.scene, lights, wl.models)
    for count in (2, 4, 7):
        wl.camera.max_light_count = count
        print("tied lights, max light count", count, flush=True)
        compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))
