"""Where the reference's GPU code departs from its CPU code (phs_cpu.cpp), which this build and
its oracle follow, and how far the benchmark processes reach those departures.

1. Breit-Wigner maps (MAP_SCHANNEL) whose resonance lies outside the branch's kinematic range,
   M^2 outside [msq_min, msq_max]: the GPU code falls back to the linear map
   (mappings_gpu.hpp:26-27); the CPU code keeps the Breit-Wigner map, whose constants then cover
   a tail of the resonance (phs_cpu.cpp:21-30, 452-462).  The two give different msq for the same
   random, so at those branches the reference's GPU and CPU paths disagree; this build agrees with
   the CPU path (the golden vectors of the 2->6 trees pin it).
2. The GPU ok flag also requires msq0 >= 0 (phs.cu:603); the CPU one does not (phs_cpu.cpp:225).
   The extra term changes no result: a negative msq0 makes m0 = sqrt(msq0) NaN, and m0 > m1 + m2
   is then false on both paths.
"""
import numpy as np

import oracle
from gpu_phs_amd import workloads as W
from gpu_phs_amd.process import MAP_SCHANNEL


def bw_outside_range(proc):
    """(channel, bincode, M, m_min, m_max) of every Breit-Wigner branch with M^2 outside
    [msq_min, msq_max] (m_max = sqrt(s) - m_tot + m_min, phs_cpu.cpp:458)."""
    ms = oracle.mass_sums(proc)
    root = proc.n_prt_out - 1
    out = []
    for c in range(proc.n_channels):
        for b in range(proc.n_prt_out - 1):
            if proc.has_children[c, b] and proc.map_ids[c, b] == MAP_SCHANNEL:
                m = proc.map_masses[c, b]
                m_max = proc.sqrts - ms[c, root] + ms[c, b]
                if not (ms[c, b] ** 2 <= m * m <= m_max ** 2):
                    out.append((c, b + 1, m, ms[c, b], m_max))
    return out


def test_breit_wigner_fork_reach_of_the_benchmark_processes():
    for proc in (W.process_c0(), W.process_c1(), W.process_c2()):
        assert bw_outside_range(proc) == []
    rows = bw_outside_range(W.process_2to6())
    print("2->6 Breit-Wigner branches below their kinematic range:", rows)
    # six branches of the sixteen 2->6 trees: M = 420 GeV on branches of three or more 173 GeV
    # particles (m_min >= 519 GeV)
    assert [(c, b) for c, b, *_ in rows] == [(0, 62), (4, 35), (5, 15), (7, 27), (11, 53), (14, 53)]
    assert all(m == 420.0 and m_min >= 519.0 for _, _, m, m_min, _ in rows)


def test_breit_wigner_fork_changes_msq():
    """At such a branch the CPU map (followed here) and the GPU reference's linear fallback give
    different msq for one random: the fork is visible at any tolerance."""
    c, b, m, m_min, m_max = bw_outside_range(W.process_2to6())[1]
    w = W.process_2to6().map_widths[c, b - 1]
    s = 3000.0 ** 2
    a, _ = oracle.constants(W.process_2to6())
    for x in (0.1, 0.5, 0.9):
        msq_cpu, _ = oracle.map_msq(MAP_SCHANNEL, x, s, m, w, m_min ** 2, m_max ** 2, a[c, b - 1])
        msq_lin = (1 - x) * m_min ** 2 + x * m_max ** 2
        assert m_min ** 2 <= msq_cpu <= m_max ** 2
        assert abs(msq_cpu - msq_lin) > 1e-3 * msq_lin


def test_negative_msq_fails_the_cpu_cut_too():
    """phs.cu:603's msq0 >= 0: with msq0 < 0 the CPU condition of phs_cpu.cpp:225 is already
    false, because every comparison with m0 = NaN is."""
    msq0, msq1, msq2, m_max = -1e-9, 0.0, 0.0, 100.0
    with np.errstate(invalid="ignore"):
        m0, m1, m2 = np.sqrt(np.float64(msq0)), 0.0, 0.0
    lda = (msq0 - msq1 - msq2) ** 2 - 4 * msq1 * msq2
    assert np.isnan(m0)
    assert not (lda > 0 and m0 > m1 + m2 and m0 <= m_max)
