"""CPU-only checks of the boundary: the C-ABI library loads and exports every
function include/scafes_hip.h declares, host-side argument checks refuse
bad input without touching a GPU, the checkpoint schedule is binomial,
and the C++ host API's index math matches the reference's numbering."""
import os
import re
import subprocess

import pytest

from masterthesis_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "scafes_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(scafes_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_side_argument_checks_need_no_gpu():
    L = capi.load()
    assert L.scafes_heat7var_num_cfgs() >= 10
    tx = capi.C.c_int()
    assert L.scafes_heat7var_cfg_desc(999, capi.C.byref(tx), None, None, None, None) == -1
    assert b"bad cfg" in L.scafes_last_error()
    # one-step tiles first, then the two-, three- and four-steps-per-sweep tiles
    st = capi.C.c_int()
    spans = []
    for cfg in range(L.scafes_heat7var_num_cfgs()):
        assert L.scafes_heat7var_cfg_desc(cfg, None, None, None, None, capi.C.byref(st)) == 0
        spans.append(st.value)
    assert set(spans) == {1, 2, 3, 4} and spans == sorted(spans)
    # null context is rejected before any HIP call
    assert L.scafes_fill(None, None, capi.F32, 10, 0.0) == -1
    n = capi.C.c_int(-7)
    assert L.scafes_batch2d_lds_limit(capi.C.byref(n)) == 0 and n.value == 256


def binom(a, b):
    from math import comb
    return comb(a, b)


@pytest.mark.parametrize("N", [1, 2, 7, 30, 100, 2000])
def test_checkpoint_schedule_is_binomial(N):
    L = capi.load()
    f = L.scafes_checkpoint_forward_steps
    assert f(N, N) == N  # store-all: every state computed once
    assert f(N, 0) == (N * (N - 1)) // 2 + 1  # no snapshot: recompute from T^0 each time
    prev = None
    for s in [0, 1, 2, 3, 5, 10, 30, 62, 200]:
        v = f(N, s)
        if prev is not None:
            assert v <= prev
        prev = v
        if 1 <= s < N:
            # Griewank's bound: with r the repetition number, at most r*N
            r = 0
            while binom(s + r, s) < N:
                r += 1
            assert v <= r * N + 1


def test_host_api_unit_program():
    exe = os.path.join(ROOT, "masterthesis_amd", "bin", "host_api_unit")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "masterthesis_amd", "host")])
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failed" in p.stdout


def test_parameters_reject_multi_partition():
    exe = os.path.join(ROOT, "masterthesis_amd", "bin", "heat_timelin_3d")
    p = subprocess.run([exe, "--nPartitions=2x1x1"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "one partition" in p.stderr


def test_compute_gradients_option_is_refused():
    """--computeGradients (ADOL-C tapes in the reference) is refused with a
    message naming the C-ABI field gradients, not silently ignored."""
    exe = os.path.join(ROOT, "masterthesis_amd", "bin", "heat_timelin_3d")
    p = subprocess.run([exe, "--computeGradients=1"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "computeGradients" in p.stderr and "scafes_heat7var_gradient" in p.stderr


def test_problem_without_device_kernel_reports_it():
    """A problem class with no DeviceTraits specialisation compiles against
    the host API unchanged, and iterateOverTime() reports that it has no
    device update instead of running the hooks on the CPU."""
    src = r'''
#include "ScaFES.hpp"
#include <iostream>
template <typename CT, std::size_t DIM>
class Plain : public ScaFES::Problem<Plain<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<Plain<CT, DIM>, CT, DIM> Base;
    using Base::Base;
    void evalInner(std::vector<ScaFES::DataField<CT, DIM>>&, ScaFES::Ntuple<int, DIM> const&, int const&) {}
    void evalBorder(std::vector<ScaFES::DataField<CT, DIM>>&, ScaFES::Ntuple<int, DIM> const&, int const&) {}
    template <typename TT> void initInner(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const&,
                                         ScaFES::Ntuple<int, DIM> const& x, int const&) { v[0](x) = 1.0; }
    template <typename TT> void initBorder(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const&,
                                          ScaFES::Ntuple<int, DIM> const& x, int const&) { v[0](x) = 0.0; }
    template <typename TT> void updateInner(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                                           ScaFES::Ntuple<int, DIM> const&, int const&) {}
    template <typename TT> void updateBorder(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                                            ScaFES::Ntuple<int, DIM> const&, int const&) {}
    template <typename TT> void updateInner2(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                                            ScaFES::Ntuple<int, DIM> const&, int const&) {}
    template <typename TT> void updateBorder2(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                                             ScaFES::Ntuple<int, DIM> const&, int const&) {}
};
int main(int argc, char* argv[]) {
    ScaFES::Parameters params(argc, argv);
    ScaFES::GridGlobal<3> gg(params);
    Plain<double, 3> p(params, gg, false, {"U"}, {1}, {false});
    try { p.iterateOverTime(); } catch (std::runtime_error const& e) { std::cout << "ERR " << e.what() << std::endl; return 3; }
    return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "plain.cpp")
        exe = os.path.join(d, "plain")
        open(cpp, "w").write(src)
        host = os.path.join(ROOT, "masterthesis_amd", "host")
        subprocess.check_call(["g++", "-std=c++11", "-O1", "-I", os.path.join(host, "ScaFES"), "-I",
                               os.path.join(ROOT, "include"), cpp, "-o", exe, "-L",
                               os.path.join(ROOT, "masterthesis_amd", "lib"), "-lscafes_hip",
                               "-Wl,-rpath," + os.path.join(ROOT, "masterthesis_amd", "lib")])
        p = subprocess.run([exe, "--nNodes=8x8x8"], capture_output=True, text=True, timeout=60)
        assert p.returncode == 3
        assert "no CPU update path" in p.stdout
