"""CPU checks of the C ABI (include/hacc_sr.h): the built library loads and
exports every declared entry point, the ctypes mirrors of the public structs
have the C layout, and argument validation runs before any device work.
No compute call is made (there is no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from haccsr import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hacc_sr.h")
LIB = os.path.join(ROOT, "hacc-coral_amd", "libhaccsr.so")
LIBCXX = os.path.join(ROOT, "hacc-coral_amd", "libhaccsr_cxx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("hsr_create", "hsr_destroy", "hsr_build", "hsr_lists", "hsr_forces",
                 "hsr_kick", "hsr_adjoint", "hsr_build_many", "hsr_lists_many", "hsr_drift",
                 "hsr_resort", "hsr_map2", "hsr_subcycle"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with make -C hacc-coral_amd"
    L = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_cxx_dropin_library_loads():
    assert os.path.exists(LIBCXX)
    C.CDLL(LIBCXX)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "hacc_sr.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(hsr_params), sizeof(hsr_node),
         sizeof(hsr_stats), sizeof(hsr_tuning));
  printf("%zu %zu %zu %zu\n", offsetof(hsr_params, nd), offsetof(hsr_params, fsrrmax),
         offsetof(hsr_params, ppc), offsetof(hsr_node, xc));
  printf("%zu %zu\n", offsetof(hsr_stats, build_ms), offsetof(hsr_tuning, use_graph));
  printf("%zu %zu %zu %zu %zu\n", sizeof(hsr_subcycle_params), offsetof(hsr_subcycle_params, ng),
         offsetof(hsr_subcycle_params, pp), offsetof(hsr_subcycle_params, fscal),
         offsetof(hsr_stats, kick_coeff));
  printf("%zu %zu %zu %zu\n", sizeof(hsr_domain), offsetof(hsr_domain, ng_local_total),
         offsetof(hsr_domain, grid2phys_vel), offsetof(hsr_stats, refresh_dead));
  printf("%zu %zu %zu\n", offsetof(hsr_stats, unmasked_pairs), offsetof(hsr_tuning, unroll),
         offsetof(hsr_params, n_interp));
  printf("%zu %zu\n", offsetof(hsr_stats, graph_captures), offsetof(hsr_stats, graph_replays));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    got = [int(v) for v in out]
    want = [C.sizeof(_lib.Params), C.sizeof(_lib.Node), C.sizeof(_lib.Stats),
            C.sizeof(_lib.Tuning), _lib.Params.nd.offset, _lib.Params.fsrrmax.offset,
            _lib.Params.ppc.offset, _lib.Node.xc.offset, _lib.Stats.build_ms.offset,
            _lib.Tuning.use_graph.offset, C.sizeof(_lib.SubcycleParams),
            _lib.SubcycleParams.ng.offset, _lib.SubcycleParams.pp.offset,
            _lib.SubcycleParams.fscal.offset, _lib.Stats.kick_coeff.offset,
            C.sizeof(_lib.Domain), _lib.Domain.ng_local_total.offset,
            _lib.Domain.grid2phys_vel.offset, _lib.Stats.refresh_dead.offset,
            _lib.Stats.unmasked_pairs.offset, _lib.Tuning.unroll.offset,
            _lib.Params.n_interp.offset, _lib.Stats.graph_captures.offset,
            _lib.Stats.graph_replays.offset]
    assert got == want
    assert _lib.NODE_DTYPE.itemsize == C.sizeof(_lib.Node)


def test_create_validates_arguments_before_device_work():
    L = _lib.load()
    h = C.c_void_p()
    assert L.hsr_create(None, None) == _lib.HSR_ERR_ARG
    p = _lib.Params()
    p.tdpts, p.nd, p.fsrrmax, p.rsm, p.ppc = 2, 128, 3.12, 0.1, 0.9
    assert L.hsr_create(C.byref(h), C.byref(p)) == _lib.HSR_ERR_UNSUPPORTED
    assert b"TDPTS" in L.hsr_last_error()
    p.tdpts, p.nd = 1, 0
    assert L.hsr_create(C.byref(h), C.byref(p)) == _lib.HSR_ERR_ARG
    p.nd, p.fsrrmax = 128, 0.0
    assert L.hsr_create(C.byref(h), C.byref(p)) == _lib.HSR_ERR_ARG
    assert L.hsr_version().startswith(b"hacc-sr")


def test_no_device_is_an_error_not_a_fallback():
    """Without a HIP device the product path raises; there is no CPU path."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    with pytest.raises(_lib.HsrError, match="no HIP device"):
        _lib.Context()


def test_domain_init_matches_the_oracle_and_validates():
    """hsr_domain_init (host arithmetic, no device) derives the one-rank
    Domain exactly as the oracle's restatement of Domain::initialize."""
    from oracle import oracle as O
    fields = ["ng_overload", "ng_local_total", "oL", "phys2grid_pos", "grid2phys_pos",
              "phys2grid_vel", "grid2phys_vel"]
    for rL, oL, ng in [(256, 3.2, 256), (1000, 10, 512), (100, 8, 64), (64, 0, 64),
                       (3.7, 0.3, 17), (512, 12.5, 1024), (1.0, 1.0, 3)]:
        a, b = _lib.domain(rL, oL, ng), O.domain(rL, oL, ng)
        for f in fields:
            assert getattr(a, f) == getattr(b, f), (rL, oL, ng, f)
    L = _lib.load()
    for rL, oL, ng in [(0, 1, 64), (64, -1, 64), (64, 65, 64), (64, 1, 0),
                       (float("nan"), 1, 64), (64, float("nan"), 64), (2000, 2, 2000)]:
        d = _lib.Domain(rL=rL, oL_in=oL, ng=ng)
        assert L.hsr_domain_init(C.byref(d)) == _lib.HSR_ERR_ARG, (rL, oL, ng)
