"""Host-side parts of the EMFDTD C ABI (no GPU needed): the CPML tables,
the per-node ga / gb coefficients, dt and the source are computed on the
host in the reference's long double / double expressions and must equal the
oracle's bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from masterthesis_amd import capi


@pytest.fixture(scope="module")
def L():
    return capi.load()


@pytest.mark.parametrize("d,m_a,m,k_max,s_max,a_max", [(5, 1.0, 3.0, 1.0, 0.01222, 1e-3),
                                                       (10, 1.0, 3.0, 1.0, 0.01222, 1e-3),
                                                       (8, 2.0, 4.0, 5.0, 0.02, 0.05),
                                                       (6, 1.0, 3.0, 1.0, 0.0, 0.0)])
def test_tables_match_oracle(L, d, m_a, m, k_max, s_max, a_max):
    dt = L.scafes_emfdtd_dt(0.5)
    assert dt == O.emfdtd_dt(0.5)
    p = capi.emfdtd_params(16, d=d, m_a=m_a, m=m, k_max=k_max, s_max=s_max, a_max=a_max)
    want = O.cpml_tables(d, m_a=m_a, m=m, k_max=k_max, s_max=s_max, a_max=a_max, dt=dt)
    for got, ref in zip((p.b, p.c, p.ook), want):
        assert np.array_equal(np.array(got[:d + 1]), ref)


def test_ga_gb_and_source_match_oracle(L):
    dt = O.emfdtd_dt(0.5)
    sigma = O.ellipsoid_sigma(20, (10, 9, 11), (6, 6, 6)).ravel()
    sigma[:50] = np.geomspace(1e-300, 1e12, 50)
    ga, gb = capi.emfdtd_ga_gb(sigma, dt)
    rga, rgb = O.ga_gb(sigma, dt)
    assert np.array_equal(ga, rga) and np.array_equal(gb, rgb)
    for t in (0.0, 3.0, 39.0, 40.0, 41.0, 100.0):
        for fm in (0.0, 3e7):
            assert L.scafes_emfdtd_source(t, dt, fm, 40.0, 8.0) == O.mod_gauss(t, dt, fm, 40.0, 8.0)


def test_bad_tables_arguments(L):
    b = (C.c_double * 70)()
    dp = C.POINTER(C.c_double)
    assert L.scafes_emfdtd_tables(1, 1.0, 3.0, 1.0, 0.01, 1e-3, 1e-9, C.cast(b, dp), C.cast(b, dp), C.cast(b, dp)) != 0
    assert L.scafes_emfdtd_tables(65, 1.0, 3.0, 1.0, 0.01, 1e-3, 1e-9, C.cast(b, dp), C.cast(b, dp), C.cast(b, dp)) != 0
