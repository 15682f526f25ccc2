"""The C-ABI library loads, exports every entry point include/t3d.h declares,
and agrees with the ctypes mirror on struct layouts (no GPU needed)."""
import ctypes as C

from tetra3d_amd import _abi as A
from tetra3d_amd import runtime


def test_library_exports_every_declared_symbol():
    L = runtime.lib()
    syms = runtime.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_abi_version():
    assert runtime.lib().t3d_abi_version() == A.T3D_ABI_VERSION


def test_struct_sizes_match_ctypes():
    sizes = runtime.struct_sizes()
    from tetra3d_amd import serving as S
    mirror = [A.Vertex, A.MeshDesc, A.MaterialDesc, A.LightDesc, A.ModelDesc, A.CameraDesc, A.WorldDesc,
              A.RenderDesc, A.PairResult, A.KernelTime, A.Tuning, S.ServerConfig, S.JobInfo, S.ServerStats,
              A.BonePose, A.AnimTrack, A.AnimChannel]
    assert sizes == [C.sizeof(s) for s in mirror]
    assert sizes[0] == 48  # ebiten.Vertex: 12 float32


def test_context_without_device_fails_loudly():
    # On a machine without a HIP device the context must refuse, never fall back.
    import pytest
    L = runtime.lib()
    ctx = C.c_void_p()
    rc = L.t3d_ctx_create(0, 0, C.byref(ctx))
    if rc == 0:
        L.t3d_ctx_destroy(ctx)
        pytest.skip("a HIP device is present")
    assert rc in (A.T3D_ERR_DEVICE, A.T3D_ERR_INVALID_ARGUMENT)
