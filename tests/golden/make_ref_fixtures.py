"""Writes tests/golden/ref_fixtures.npz: golden vectors produced by the compiled
reference (oracle/_ref/libmsref.so, built from the MSModel sources by
oracle/ref.mk, LAPACK = conda MKL).

Every array is an input or an output of a reference function on a stated
input; nothing of the reference's source is stored.  Run from the repository
root in a container that holds the reference tree:

    make -f oracle/ref.mk && python tests/golden/make_ref_fixtures.py

Medium convention for a stated radius R (the reference only knows unit
density): positions = fill_medium(seed) * (R / radius_of_ball), one FP64
multiply per coordinate after the reference's own fill.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyref as R  # noqa: E402

OUT = os.path.join(HERE, "ref_fixtures.npz")
out = {}


def put(name, arr):
    out[name] = np.asarray(arr)


def medium(d, n, seed, model=0, p=(0.1,), shape=1, ratio=1.0, radius=None):
    m = R.RefMedium(d, n, model, p, shape, ratio).fill(seed)
    if radius is not None:
        pos = m.pos * (radius / R.lib().radius_of_ball(m.med))
        m.set_pos(pos)
    return m


# ----------------------------------------------------------------- RNG streams
seeds = np.array([0, 1, 2, 3, 100, 12345, 2 ** 40 + 7, 2 ** 64 - 1], dtype=np.uint64)
put("rng_seeds", seeds)
put("rng_streams", np.stack([R.rng_stream(int(s), 32) for s in seeds]))

# ----------------------------------------------------------------- positions
media = [  # d, n, shape, ratio, seed
    (3, 200, 1, 1.0, 1), (3, 1000, 1, 1.0, 2), (3, 500, 1, 1.0, 3), (3, 1000, 1, 1.0, 3),
    (2, 50, 1, 1.0, 1), (2, 20, 1, 1.0, 1), (1, 40, 1, 1.0, 1), (4, 120, 1, 1.0, 4),
    (3, 64, 0, 1.0, 7), (3, 64, 3, 1.0, 8), (2, 49, 2, 1.0, 1), (3, 100, 1, 2.5, 9),
    (2, 1000, 1, 1.0, 100), (2, 300, 1, 1.0, 101),
]
put("media_spec", np.array(media, dtype=np.float64))
for i, (d, n, shape, ratio, seed) in enumerate(media):
    put("media_pos_%d" % i, medium(d, n, seed, shape=shape, ratio=ratio).pos)

# ----------------------------------------------------------------- scalar Green functions
rng = np.random.default_rng(7)
ks = np.concatenate([rng.uniform(0.2, 40, 24) + 1j * rng.uniform(-2.5, 0.3, 24),
                     np.array([10.0, 1e-3 + 0j, 5 - 2j, 0.5 - 0.01j, 24.048 - 0.2j, -3 - 1j])])
rs = np.concatenate([rng.uniform(0.01, 20, 24), np.array([1.0, 0.5, 3.0, 0.05, 2.0, 1.0])])
put("green_k", ks)
put("green_r", rs)
for d in (1, 2, 3, 4):
    put("green_d%d" % d, [R.green(d, k, r) for k, r in zip(ks, rs)])
    put("green_imag_d%d" % d, [R.green_imag(d, k, r) for k, r in zip(ks, rs)])
    put("dk_green_d%d" % d, [R.dk_green(d, k, r) for k, r in zip(ks, rs)])

# ----------------------------------------------------------------- invf per model
models = [(0, (0.1,)), (0, (0.4,)), (0, (0.05,)), (1, (0.2,)), (2, (5.0, -0.5)), (3, ())]
invf_k = np.array([0.5 - 0.1j, 1.0, 3.0 - 0.5j, 7.5 - 1j, 10.0, 15.0 - 2j, 24.0 - 0.3j, 30 - 1.5j,
                   12.0 + 0.2j, 2.0 - 0.01j, 40.0 - 1.0j, 5.0])
put("invf_k", invf_k)
for mi, (model, p) in enumerate(models):
    put("invf_model_%d" % mi, np.array([model] + list(p) + [0.0] * (2 - len(p)), dtype=np.float64))
    for d in (1, 2, 3, 4):
        m = R.RefMedium(d, 2, model, p if p else (0.0,))
        put("invf_%d_d%d" % (mi, d), [m.invf(k) for k in invf_k])

# ----------------------------------------------------------------- full matrices (small N)
for d in (1, 2, 3, 4):
    m = medium(d, 24, 11 + d, p=(0.1,))
    put("mat_pos_d%d" % d, m.pos)
    put("mat_ms_d%d" % d, m.matrix(7.0 - 0.3j))
    put("mat_dk_d%d" % d, m.matrix(7.0 - 0.3j, "dk"))
    put("mat_norm_d%d" % d, m.matrix(7.0 - 0.3j, "normalized"))

# ----------------------------------------------------------------- tracelog
# (a) the reference test medium: d=2 N=50 hardsphere 0.4 seed 1 (test_ms_matrix.c:36-57)
m = medium(2, 50, 1, p=(0.4,))
tk = np.array([1 - 1j, 3 - 2j, 6 - 6j, 10 - 0.5j, 10 - 10j])
put("tl_test_k", tk)
put("tl_test_val", [R.tracelog(m.matrix(k)) for k in tk])
# (b) config-1 medium (d=3, N=200, unit ball, seed 1, alpha=0.1) on a 16x8 subgrid of the map
m = medium(3, 200, 1, p=(0.1,), radius=1.0)
put("tl_cfg1_pos", m.pos)
xs = np.linspace(1, 20, 16)
ys = np.linspace(0, -1, 8)
kk = (xs[None, :] + 1j * ys[:, None]).ravel()
put("tl_cfg1_k", kk)
put("tl_cfg1_val", [R.tracelog(m.matrix(k)) for k in kk])
# (c) config-3 medium (d=3, N=500, unit ball, seed 3, alpha=0.05), 8 k
m = medium(3, 500, 3, p=(0.05,), radius=1.0)
put("tl_cfg3_pos", m.pos)
k3 = np.array([5 + 0j, 8 - 0.5j, 12.3 - 1.1j, 17 - 2j, 21.5 - 0.25j, 25 - 1.7j, 29.6 - 0.03j, 10 - 1.9j])
put("tl_cfg3_k", k3)
put("tl_cfg3_val", [R.tracelog(m.matrix(k)) for k in k3])
# (d) headline medium: d=3, N=1000, unit ball, seed 3, alpha=0.05, 4 k
m = medium(3, 1000, 3, p=(0.05,), radius=1.0)
put("tl_n1000_pos", m.pos)
k4 = np.array([5 + 0j, 11.5 - 0.7j, 19.9 - 1.3j, 28.2 - 2j])
put("tl_n1000_k", k4)
put("tl_n1000_val", [R.tracelog(m.matrix(k)) for k in k4])
# (e) d=2 Hankel medium: N=400 unit disk seed 100, alpha=0.1, 4 k
m = medium(2, 400, 100, p=(0.1,), radius=1.0)
put("tl_d2_pos", m.pos)
k5 = np.array([1.5 - 0.1j, 9.0 - 0.5j, 20.0 - 0.9j, 37.0 - 0.2j])
put("tl_d2_k", k5)
put("tl_d2_val", [R.tracelog(m.matrix(k)) for k in k5])

# ----------------------------------------------------------------- dk_log_det
m = medium(2, 50, 1, p=(0.4,))
dk_k = np.array([1 - 1j, 10 - 0.5j, 10 - 10j, 0.1, 1., 3., 10., 100.])
put("dk_test_k", dk_k)
put("dk_test_val", [m.dk_log_det(k) for k in dk_k])
m = medium(3, 200, 1, p=(0.1,), radius=1.0)
dk3 = np.array([2 - 0.2j, 7.5 - 0.5j, 13 - 0.9j, 19 - 0.1j])
put("dk_cfg1_k", dk3)
put("dk_cfg1_val", [m.dk_log_det(k) for k in dk3])
m = medium(2, 200, 101, p=(0.1,), radius=1.0)
put("dk_d2_pos", m.pos)
dk2 = np.array([3 - 0.3j, 15 - 0.6j, 33 - 0.1j])
put("dk_d2_k", dk2)
put("dk_d2_val", [m.dk_log_det(k) for k in dk2])

# ----------------------------------------------------------------- solves and wavefunctions
# (a) test_ms_matrix.c:209-246: d=2 N=20 hardsphere 0.4 seed 1, k=1, plane wave theta=0
m = medium(2, 20, 1, p=(0.4,))
phi = m.plane_wave(1.0, 0.0)
a, _, _, info = R.zgesv(m.matrix(1.0), phi)
put("sol_test_a", a)
put("sol_test_xs", [m.total_cross_section(k) for k in (0.1, 1., 3., 10., 100.)])
# (b) config-2 medium scaled down: d=3, N=300, radius 5*(300/1000)^(1/3), seed 2, alpha=0.1, k=10
rad = 5.0 * (300 / 1000) ** (1 / 3)
m = medium(3, 300, 2, p=(0.1,), radius=rad)
put("sol_cfg2_pos", m.pos)
phi = np.exp(1j * 10.0 * m.pos[:, 0])
a, _, _, info = R.zgesv(m.matrix(10.0), phi)
put("sol_cfg2_a", a)
put("sol_cfg2_psi", m.wave_grid(a, 10.0, 17, 13, -8, 8, -6, 6))
# (c) d=2, N=150 unit-density disk, seed 1, maximum model, k=10, 11x9 grid (examples/wavefun.conf shape)
m = medium(2, 150, 1, model=3, p=(0.0,))
put("sol_d2_pos", m.pos)
phi = np.exp(1j * 10.0 * m.pos[:, 0])
a, _, _, info = R.zgesv(m.matrix(10.0), phi)
put("sol_d2_a", a)
put("sol_d2_psi", m.wave_grid(a, 10.0, 11, 9, -8, 8, -8, 8))
# (d) spherical source, d=3 N=80 seed 5, k = 4 - 0.1i
m = medium(3, 80, 5, p=(0.1,))
put("sol_sph_pos", m.pos)
phi = np.array([R.green(3, 4 - 0.1j, np.linalg.norm(x)) for x in m.pos])
a, _, _, info = R.zgesv(m.matrix(4 - 0.1j), phi)
put("sol_sph_a", a)
put("sol_sph_psi", m.wave_grid(a, 4 - 0.1j, 9, 7, -5, 5, -4, 4, source=1))

# ----------------------------------------------------------------- traceinv / inverse iteration (3x3 test matrix)
A3 = np.array([-1.997 + 0.45j, 1.875 - 1.47j, -1.641 + 0.204j, -1.74 - 1.613j, -1.197 - 1.071j,
               1.883 - 1.765j, 1.958 + 0.195j, 1.278 + 1.105j, -1.08 + 1.842j]).reshape(3, 3, order="F")
put("a3", A3)
put("a3_traceinv", R.traceinv(A3))

# ----------------------------------------------------------------- Aberth resonances
m = medium(2, 20, 1, p=(0.4,))
dom = (1.0, 10.0, -1.0, 0.0)
starts = np.array([dom[0] + (dom[1] - dom[0]) * x + 1j * (dom[2] + (dom[3] - dom[2]) * y)
                   for x, y in zip(np.linspace(0.07, 0.93, 12), np.linspace(0.9, 0.1, 12))])
put("ab_d2_starts", starts)
put("ab_d2_dom", np.array(dom))
put("ab_d2_roots", m.aberth(starts, dom, nexpsup=4, maxit=60, toler=1e-12))
m = medium(3, 30, 2, p=(0.1,))
put("ab_d3_pos", m.pos)
dom = (2.0, 12.0, -1.5, 0.0)
starts = np.array([dom[0] + (dom[1] - dom[0]) * x + 1j * (dom[2] + (dom[3] - dom[2]) * y)
                   for x, y in zip(np.linspace(0.05, 0.95, 10), np.linspace(0.85, 0.15, 10))])
put("ab_d3_starts", starts)
put("ab_d3_dom", np.array(dom))
put("ab_d3_roots", m.aberth(starts, dom, nexpsup=4, maxit=60, toler=1e-12))

# ----------------------------------------------------------------- k-point generators (bit-exact)
# get_arg_cmap (complex_map_util.c:27-31), get_arg_ccut (complex_cut_util.c:32-36),
# init_logcut + get_arg_lcut (log_cut_util.c:72-101)
grid_map = [(13, 7, 1.0, 20.0, -1.0, 0.0), (64, 64, 5.0, 30.0, -2.0, 0.0), (5, 3, -2.5, 7.25, -3.0, 0.5)]
put("grid_map_spec", np.array(grid_map, dtype=np.float64))
for i, g in enumerate(grid_map):
    put("grid_map_%d" % i, R.kgrid_map(*g))
grid_cut = [(4, 1 - 0.5j, 5 - 0.1j), (7, 2 + 0j, 12 - 1j), (20, 3 - 0.01j, 3 - 2j), (33, 0.5 - 1j, 40 - 0.2j)]
put("grid_cut_ns", np.array([g[0] for g in grid_cut], dtype=np.int64))
put("grid_cut_ends", np.array([[g[1], g[2]] for g in grid_cut], dtype=np.complex128))
for i, g in enumerate(grid_cut):
    put("grid_cut_%d" % i, R.kgrid_cut(*g))
grid_log = [(37, 12.5, -1e-4, -2.0), (10, 5.0, -1.0, -1e-3), (2, 30.0, -0.5, -0.25)]
put("grid_log_spec", np.array(grid_log, dtype=np.float64))
for i, g in enumerate(grid_log):
    put("grid_log_%d" % i, R.kgrid_logcut(g[0], *g[1:]))

# ----------------------------------------------------------------- inverse iteration (3x3 test matrix)
conv, v, lam = R.inverse_iteration(A3, np.ones(3) / np.sqrt(3), 30, 1e-12)
put("a3_ivit_vec", v)
put("a3_ivit_lambda", lam)

# ----------------------------------------------------------------- [eigenstate_list] driver, all methods
# run_eigenstate_list_serial (ms_core.c:406-430): Halton starts sorted by -Im k, the root
# finder of the method (0 determinant/Aberth, 1 mineigval/Muller, 2 invtraceinv/Muller),
# then compute_eigenstates: per state the root, mu and the eigenvector (50 inverse
# iterations, 1e-10, from the constant unit vector).
eig_cases = [  # tag, d, n, seed, alpha, dom, ntarget, maxit, toler, method
    ("det_d2", 2, 20, 1, 0.4, (1.0, 10.0, -1.0, 0.0), 12, 60, 1e-12, 0),
    ("mev_d2", 2, 20, 1, 0.4, (1.0, 10.0, -1.0, 0.0), 12, 30, 1e-12, 1),
    ("iti_d2", 2, 20, 1, 0.4, (1.0, 10.0, -1.0, 0.0), 12, 30, 1e-12, 2),
    ("mev_d3", 3, 30, 2, 0.1, (2.0, 12.0, -1.5, 0.0), 10, 30, 1e-12, 1),
    ("iti_d3", 3, 30, 2, 0.1, (2.0, 12.0, -1.5, 0.0), 10, 30, 1e-12, 2),
    ("mev_d2n60", 2, 60, 3, 0.2, (2.0, 9.0, -0.8, 0.0), 16, 30, 1e-12, 1),
    ("iti_d2n60", 2, 60, 3, 0.2, (2.0, 9.0, -0.8, 0.0), 16, 30, 1e-12, 2),
]
for tag, d, n, seed, alpha, dom, nt, maxit, toler, method in eig_cases:
    m = medium(d, n, seed, p=(alpha,))
    kr, mu, ev = m.eigenstate_list(method, nt, dom, maxit, toler)
    put("eig_%s_spec" % tag, np.array([d, n, seed, alpha, dom[0], dom[1], dom[2], dom[3], nt, maxit, toler, method]))
    put("eig_%s_pos" % tag, m.pos)
    put("eig_%s_kroot" % tag, kr)
    put("eig_%s_mu" % tag, mu)
    put("eig_%s_eigv" % tag, ev)

np.savez_compressed(OUT, **out)
print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(out), "arrays")
