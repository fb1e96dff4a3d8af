"""Krylov solves: Solver(linear_solver='cg' | 'gmres' | 'bcgs' | 'bicgstab',
preconditioner='jacobi' | 'none'), as the reference's demos ask
(demo/laplace_ffc.py:151, weak_bcs_ffc.py:182) and as PETSc 3.3's KSPCG,
KSPGMRES and KSPBCGS with left PCJACOBI or PCNONE run them
(pyop2/petsc_base.py:172-220; cusp names, pyop2/cuda.py:504-508).

CPU: the oracle (oracle_gmres, oracle_bcgs, oracle_pcg) is pinned by
  * scipy's independent GMRES, BiCGStab and CG on the explicitly
    preconditioned system D^-1 A x = D^-1 b (CG: scipy's PCG with M = D^-1),
    and unpreconditioned: the same iteration counts and residual histories
    (GMRES estimates, BiCGStab and CG true residuals);
  * brute force: the GMRES residual after k steps is the minimum of
    |D^-1 (b - A x)| over x in the k-th Krylov space, computed by least
    squares on a twice-orthogonalised basis, in the first cycle and, from
    the restarted iterate, in the second;
  * closed cases: zero right-hand side (ATOL at 0 iterations), a diagonal
    matrix with power-of-two entries (BiCGStab's t = 0 exit gives x = b/d
    exactly at 1 iteration), max_it and dtol exits, the dense solution.
GPU: the C-ABI solve against the oracle on the same systems (reason equal,
iterations equal or one apart, as the last residual may cross the tolerance
a step later under another summation order; residual histories within 1e-5
relative for GMRES, 1e-4 above 1e-6 of the first residual for BiCGStab;
solutions within 1e-8 of the oracle's at rtol 1e-12, as both meet the same
residual bound), with and without the preconditioner; op2-level solves of
an assembled nonsymmetric P1 convection-diffusion operator with Dirichlet
rows (zero_rows) and of SPD operators (CG) against the oracle backend (same
reason and solution, histories alike; iteration counts within a tenth for
BiCGStab on the convection-dominated case); repeated solves replaying the
library's cached graphs; at 1M rows the true residual equals the reported
one."""

import ctypes

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as sl

from oracle import sequential as seq

METHODS = ('gmres', 'bcgs')


def convdiff(nx, bx=30.0, by=15.0, seed=0, shift=0.0):
    """Upwind 5-point convection-diffusion on an nx x nx grid (plus a
    reaction term ``shift`` times the diffusion diagonal), rows scaled by
    seeded factors in [0.5, 2] (nonsymmetric, Jacobi not trivial)."""
    h = 1.0 / (nx + 1)
    n = nx * nx
    i, j = np.divmod(np.arange(n), nx)
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [np.full(n, (4 + 4 * shift) / h ** 2 + bx / h + by / h)]
    for di, dj, v in ((-1, 0, -1 / h ** 2 - bx / h), (1, 0, -1 / h ** 2),
                      (0, -1, -1 / h ** 2 - by / h), (0, 1, -1 / h ** 2)):
        ok = (i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < nx)
        rows.append(np.arange(n)[ok])
        cols.append(((i + di) * nx + j + dj)[ok])
        vals.append(np.full(ok.sum(), v))
    A = sps.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(n, n))
    s = np.random.default_rng(seed).uniform(0.5, 2.0, n)
    A = (sps.diags(s) @ A).tocsr()
    A.sort_indices()
    return A


def ora(method, A, b, rtol=1e-10, atol=1e-50, dtol=1e4, maxit=1000, restart=30, jacobi=1):
    """The oracle's solve: (x, its, rnorm, reason, history)."""
    rp = A.indptr.astype(np.int32)
    ci = A.indices.astype(np.int32)
    v = A.data.astype(np.float64)
    n = A.shape[0]
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    hist = np.zeros(maxit + 2)
    its, rn = ctypes.c_int(), ctypes.c_double()
    f = getattr(seq.rt(), 'oracle_' + method)
    f.restype = ctypes.c_int
    extra = [restart] if method == 'gmres' else []
    f.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int] * len(extra) + \
        [ctypes.c_double] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int]
    reason = f(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, bb.ctypes.data, x.ctypes.data,
               *extra, rtol, atol, dtol, maxit, ctypes.byref(its), ctypes.byref(rn),
               hist.ctypes.data, jacobi)
    return x, its.value, rn.value, reason, hist[:its.value + 1].copy()


def precond(A, b):
    d = A.diagonal()
    return (sps.diags(1.0 / d) @ A).tocsr(), b / d


def close_history(h, ref, floor):
    """Relative agreement of two residual histories where they are above
    floor (below it rounding dominates)."""
    k = min(len(h), len(ref))
    keep = ref[:k] > floor
    return np.abs(h[:k][keep] / ref[:k][keep] - 1).max()


# -- CPU: the oracle pinned ---------------------------------------------------

@pytest.mark.parametrize('restart', [30, 5])
def test_oracle_gmres_matches_scipy(restart):
    A = convdiff(12)
    b = np.random.default_rng(1).random(A.shape[0])
    x, its, rn, reason, hist = ora('gmres', A, b, rtol=1e-10, restart=restart)
    assert reason == 2
    Ah, bh = precond(A, b)
    got = []
    sl.gmres(Ah, bh, rtol=1e-10, atol=0.0, restart=restart, maxiter=500,
             callback=got.append, callback_type='pr_norm')
    got = np.array(got) * np.linalg.norm(bh)
    assert len(got) == its
    assert close_history(hist[1:], got, 1e-9 * hist[0]) < 1e-5
    assert np.abs(x - np.linalg.solve(A.toarray(), b)).max() <= 1e-8 * np.abs(x).max()


def test_oracle_bcgs_matches_scipy():
    A = convdiff(12)
    b = np.random.default_rng(1).random(A.shape[0])
    x, its, rn, reason, hist = ora('bcgs', A, b, rtol=1e-10)
    assert reason == 2
    Ah, bh = precond(A, b)
    xs = []
    sl.bicgstab(Ah, bh, rtol=1e-14, atol=0.0, maxiter=its, callback=lambda xk: xs.append(xk.copy()))
    true = np.array([np.linalg.norm(bh - Ah @ xk) for xk in xs])
    assert len(true) == its
    assert close_history(hist[1:], true, 1e-8 * hist[0]) < 1e-6
    assert np.abs(x - np.linalg.solve(A.toarray(), b)).max() <= 1e-8 * np.abs(x).max()
    # the recursive residual is the true residual
    assert abs(np.linalg.norm(bh - Ah @ x) - rn) <= 1e-12 * hist[0]


def _min_residual(Ah, r0, k):
    """min over y of |r0 - Ah Q y|, Q an orthonormal basis of the k-th
    Krylov space of Ah and r0 (Gram-Schmidt applied twice)."""
    Q = np.zeros((len(r0), k))
    q = r0 / np.linalg.norm(r0)
    for j in range(k):
        for _ in range(2):
            q = q - Q[:, :j] @ (Q[:, :j].T @ q)
        Q[:, j] = q / np.linalg.norm(q)
        q = Ah @ Q[:, j]
    y = np.linalg.lstsq(Ah @ Q, r0, rcond=None)[0]
    return np.linalg.norm(r0 - Ah @ (Q @ y))


def test_oracle_gmres_minimal_residual_brute_force():
    A = convdiff(5, seed=3)
    b = np.random.default_rng(2).random(A.shape[0])
    Ah, bh = precond(A, b)
    m = 6
    x, its, rn, reason, hist = ora('gmres', A, b, rtol=1e-13, restart=m)
    for k in range(1, m + 1):
        assert abs(hist[k] - _min_residual(Ah, bh, k)) <= 1e-9 * hist[0], k
    # the restart: x after one cycle (max_it = m), then the second cycle
    xm, itm, rnm, rsm, hm = ora('gmres', A, b, rtol=1e-13, restart=m, maxit=m)
    assert rsm == -3 and itm == m
    r1 = bh - Ah @ xm
    assert abs(hist[m] - np.linalg.norm(r1)) <= 1e-12 * hist[0]
    for k in range(1, m + 1):
        assert abs(hist[m + k] - _min_residual(Ah, r1, k)) <= 1e-9 * hist[0], k


def test_oracle_gmres_finite_termination():
    """With restart >= n, GMRES solves an n-row system in at most n steps."""
    A = convdiff(4, seed=5)
    n = A.shape[0]
    b = np.random.default_rng(4).random(n)
    x, its, rn, reason, hist = ora('gmres', A, b, rtol=1e-12, restart=n)
    assert reason in (2, 3) and its <= n
    assert np.abs(x - np.linalg.solve(A.toarray(), b)).max() <= 1e-10 * np.abs(x).max()


@pytest.mark.parametrize('method', METHODS)
def test_oracle_zero_rhs(method):
    A = convdiff(4)
    x, its, rn, reason, hist = ora(method, A, np.zeros(A.shape[0]))
    assert (its, rn, reason) == (0, 0.0, 3)
    assert not x.any()


def test_oracle_diagonal_exact():
    """D^-1 A = I with power-of-two entries: v = p = r, alpha = 1, s = 0 and
    t = 0, so BiCGStab takes its t = 0 exit with x = b/d exactly."""
    d = np.array([2.0, 4.0, 0.5, 8.0, 1.0, 0.25])
    A = sps.diags(d).tocsr()
    b = np.random.default_rng(6).random(len(d))
    x, its, rn, reason, hist = ora('bcgs', A, b)
    assert (its, rn, reason) == (1, 0.0, 2)
    assert np.array_equal(x, b / d)
    x, its, rn, reason, hist = ora('gmres', A, b)
    assert its == 1 and reason in (2, 3)
    assert np.abs(x - b / d).max() <= 1e-15 * np.abs(b / d).max()


@pytest.mark.parametrize('method', METHODS)
def test_oracle_exits(method):
    A = convdiff(8)
    b = np.random.default_rng(7).random(A.shape[0])
    x, its, rn, reason, hist = ora(method, A, b, maxit=3)
    assert (its, reason) == (3, -3)
    # dtol < 1 fails the iteration-0 test: |r0| >= dtol |r0|
    x, its, rn, reason, hist = ora(method, A, b, dtol=0.5)
    assert (its, reason) == (0, -4)
    assert not x.any()


# -- GPU: the C-ABI solve against the oracle --------------------------------------

def hip_solve(method, A, b, rtol=1e-10, atol=1e-50, dtol=1e4, maxit=1000, restart=30,
              hist=True, jacobi=1):
    import torch
    from pyop2_amd import _native as nat
    dev = torch.device('cuda', 0)
    code = {'gmres': 1, 'bcgs': 2}[method]
    n = A.shape[0]
    rp = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ci = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    v = torch.from_numpy(A.data.astype(np.float64)).to(dev)
    bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(dev)
    x = torch.full((max(n, 1),), np.nan, dtype=torch.float64, device=dev)
    nwork = nat.lib.op2hip_krylov_work(code, n, restart)
    assert nwork >= 0
    work = torch.full((max(nwork, 1),), np.nan, dtype=torch.float64, device=dev)
    h = torch.zeros(maxit + 1, dtype=torch.float64, device=dev) if hist else None
    its, rn, reason = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    s = torch.cuda.current_stream().cuda_stream
    nat.check(nat.lib.op2hip_krylov_solve(
        code, n, rp.data_ptr(), ci.data_ptr(), v.data_ptr(), bd.data_ptr(), x.data_ptr(),
        work.data_ptr(), work.numel(), restart, jacobi, rtol, atol, dtol, maxit,
        h.data_ptr() if hist else None, h.numel() if hist else 0, ctypes.byref(its),
        ctypes.byref(rn), ctypes.byref(reason), s), 'krylov')
    its = its.value
    hh = h[:its + 1].cpu().numpy() if hist else None
    return x[:n].cpu().numpy(), its, rn.value, reason.value, hh


# GMRES residuals fall monotonically and agree closely; BiCGStab's are not
# monotone and amplify rounding, so its histories are compared above 1e-6 of
# the initial residual and to 1e-4
HIST_FLOOR = {'gmres': 1e-8, 'bcgs': 1e-6, 'bicgstab': 1e-6}
HIST_TOL = {'gmres': 1e-5, 'bcgs': 1e-4, 'bicgstab': 1e-4}


def assert_parity(method, A, b, **kw):
    xo, io, rno, ro, ho = ora(method, A, b, **kw)
    xg, ig, rng, rg, hg = hip_solve(method, A, b, **kw)
    # summation order differs (workgroup partials against a sequential sum),
    # so the last residual may cross the tolerance one iteration apart
    assert rg == ro and abs(ig - io) <= 1, (method, ig, rg, io, ro)
    assert close_history(hg, ho, HIST_FLOOR[method] * ho[0]) < HIST_TOL[method]
    scale = max(np.abs(xo).max(), 1e-300)
    assert np.abs(xg - xo).max() <= 1e-8 * scale
    return xg, ig


@pytest.mark.gpu
@pytest.mark.parametrize('method', METHODS)
@pytest.mark.parametrize('nx,restart', [(12, 30), (12, 5), (40, 30), (64, 50), (40, 100)])
def test_hip_matches_oracle(method, nx, restart):
    if method == 'bcgs' and restart != 30:
        pytest.skip('restart is a GMRES parameter')
    A = convdiff(nx, seed=nx)
    b = np.random.default_rng(nx).random(A.shape[0])
    assert_parity(method, A, b, rtol=1e-12, restart=restart)


@pytest.mark.gpu
@pytest.mark.parametrize('method', METHODS)
def test_hip_exits(method):
    A = convdiff(16)
    b = np.random.default_rng(8).random(A.shape[0])
    assert_parity(method, A, b, maxit=7)
    x, its, rn, reason, h = hip_solve(method, A, b, dtol=0.5)
    assert (its, reason) == (0, -4) and not x.any()
    x, its, rn, reason, h = hip_solve(method, A, np.zeros(A.shape[0]))
    assert (its, rn, reason) == (0, 0.0, 3) and not x.any()
    x, its, rn, reason, h = hip_solve(method, A[:0, :0].tocsr(), np.zeros(0))
    assert (its, reason) == (0, 3)


@pytest.mark.gpu
def test_hip_diagonal_exact():
    d = np.array([2.0, 4.0, 0.5, 8.0, 1.0, 0.25] * 50)
    A = sps.diags(d).tocsr()
    b = np.random.default_rng(6).random(len(d))
    x, its, rn, reason, h = hip_solve('bcgs', A, b)
    assert (its, rn, reason) == (1, 0.0, 2)
    assert np.array_equal(x, b / d)
    x, its, rn, reason, h = hip_solve('gmres', A, b)
    assert its == 1 and reason in (2, 3)


@pytest.mark.gpu
@pytest.mark.parametrize('method', METHODS)
def test_hip_large_true_residual(method):
    """1M rows: the solve converges, and the reported residual is the true
    preconditioned residual (GMRES: its estimate; BiCGStab: its recursion),
    within rounding."""
    A = convdiff(1000, bx=3.0, by=1.5, seed=11, shift=0.05)
    b = np.random.default_rng(11).random(A.shape[0])
    x, its, rn, reason, h = hip_solve(method, A, b, rtol=1e-6, maxit=3000, hist=False)
    assert reason == 2, (its, reason)
    Ah, bh = precond(A, b)
    r0 = np.linalg.norm(bh)
    true = np.linalg.norm(bh - Ah @ x)
    assert true <= 1e-6 * r0 * 1.01
    assert abs(true - rn) <= 1e-9 * r0


# -- op2 level: assembled nonsymmetric operator with Dirichlet rows ---------------

ADV_DIFF = """
void adv_diff(double A[1][1], double *x[2], int i, int j)
{
  /* P1: eps grad.grad + (b . grad phi_j) phi_i with b = (3, 1), area/3 per
     test function (one-point rule of the convection term) */
  double J00 = x[1][0] - x[0][0], J01 = x[2][0] - x[0][0];
  double J10 = x[1][1] - x[0][1], J11 = x[2][1] - x[0][1];
  double det = J00 * J11 - J01 * J10;
  double area = 0.5 * fabs(det);
  double g[3][2];
  g[0][0] = (J10 - J11) / det;  g[0][1] = (J01 - J00) / det;
  g[1][0] = J11 / det;          g[1][1] = -J01 / det;
  g[2][0] = -J10 / det;         g[2][1] = J00 / det;
  double eps = 0.05;
  A[0][0] += area * (eps * (g[i][0] * g[j][0] + g[i][1] * g[j][1]))
           + (area / 3.0) * (3.0 * g[j][0] + 1.0 * g[j][1]);
}
"""


def _assembled_solve(W, mesh, f, method, rtol, restart=30):
    nodes = W.Set(mesh.nvert, 1, 'nodes')
    vnodes = W.Set(mesh.nvert, 2, 'vnodes')
    el = W.Set(mesh.nelem, 1, 'elements')
    en = W.Map(el, nodes, 3, mesh.elem_node, 'en')
    ev = W.Map(el, vnodes, 3, mesh.elem_node, 'ev')
    A = W.Mat(W.Sparsity((en, en), 'adv'), np.float64, 'A')
    x = W.Dat(vnodes, mesh.coords.copy(), np.float64, 'x')
    W.par_loop(W.Kernel(ADV_DIFF, 'adv_diff'), el(3, 3), A((en[W.i[0]], en[W.i[1]]), W.INC),
               x(ev, W.READ))
    bnd = np.where((mesh.coords[:, 0] < 1e-12) | (mesh.coords[:, 0] > 1 - 1e-12) |
                   (mesh.coords[:, 1] < 1e-12) | (mesh.coords[:, 1] > 1 - 1e-12))[0]
    A.zero_rows(bnd, 1.0)
    rhs = f.copy()
    rhs[bnd] = 0.0
    b = W.Dat(nodes, rhs, np.float64, 'b')
    u = W.Dat(nodes, np.zeros(mesh.nvert), np.float64, 'u')
    solver = W.Solver(linear_solver=method, relative_tolerance=rtol, maximum_iterations=5000,
                      gmres_restart=restart, monitor_convergence=True)
    solver.solve(A, u, b)
    return (np.array(u.data_ro).reshape(-1), solver.getIterationNumber(),
            solver.getConvergedReason(), np.array(solver.reshist))


@pytest.mark.gpu
@pytest.mark.parametrize('method', ['gmres', 'bicgstab', 'bcgs'])
def test_op2_assembled_nonsymmetric(hip, method):
    from pyop2_amd.workloads import meshes
    mesh = meshes.TriMesh(24, seed=3)
    f = np.random.default_rng(3).random(mesh.nvert)
    ug, ig, rg, hg = _assembled_solve(hip, mesh, f, method, 1e-12, restart=20)
    uo, io, ro, ho = _assembled_solve(seq, mesh, f, method, 1e-12, restart=20)
    assert rg == ro and rg > 0
    assert np.abs(ug - uo).max() <= 1e-8 * np.abs(uo).max()
    # the same method: the first iterations agree closely.  On this
    # convection-dominated operator BiCGStab's residual jumps about, and
    # rounding (summation order; matrix entries summed in another order)
    # moves its last iterations, so its count may differ by a tenth
    k = min(10, len(hg), len(ho))
    assert np.abs(hg[:k] / ho[:k] - 1).max() < 1e-6
    slack = 1 if method == 'gmres' else max(1, io // 10)
    assert abs(ig - io) <= slack, (ig, io)


@pytest.mark.gpu
@pytest.mark.parametrize('method', ['gmres', 'bcgs'])
def test_repeated_solves_replay_cached_graphs(hip, method):
    """One Solver solving one system again and again (as a time loop does)
    presents the same arguments, so the library replays its cached round
    graphs: the right-hand side's new values are used, a repeated solve is
    bit-identical, and a solve of another system in between changes
    nothing."""
    from pyop2_amd.workloads import meshes
    mesh = meshes.TriMesh(16, seed=5)
    W = hip
    nodes = W.Set(mesh.nvert, 1)
    vnodes = W.Set(mesh.nvert, 2)
    el = W.Set(mesh.nelem, 1)
    en = W.Map(el, nodes, 3, mesh.elem_node)
    ev = W.Map(el, vnodes, 3, mesh.elem_node)
    A = W.Mat(W.Sparsity((en, en)), np.float64)
    x = W.Dat(vnodes, mesh.coords.copy(), np.float64)
    W.par_loop(W.Kernel(ADV_DIFF, 'adv_diff'), el(3, 3), A((en[W.i[0]], en[W.i[1]]), W.INC),
               x(ev, W.READ))
    c = mesh.coords
    A.zero_rows(np.where((c[:, 0] < 1e-12) | (c[:, 0] > 1 - 1e-12) |
                         (c[:, 1] < 1e-12) | (c[:, 1] > 1 - 1e-12))[0], 1.0)
    rng = np.random.default_rng(5)
    b1, b2 = rng.random(mesh.nvert), rng.random(mesh.nvert)
    b = W.Dat(nodes, b1.copy(), np.float64)
    u = W.Dat(nodes, np.zeros(mesh.nvert), np.float64)
    s = W.Solver(linear_solver=method, relative_tolerance=1e-12, maximum_iterations=5000)
    s.solve(A, u, b)
    u1 = np.array(u.data_ro).ravel()
    b.data[:] = b2.reshape(-1, 1)
    s.solve(A, u, b)
    u2 = np.array(u.data_ro).ravel()
    # another system through another Solver in between
    A2 = convdiff(12, seed=2)
    hip_solve(method, A2, np.ones(A2.shape[0]))
    b.data[:] = b1.reshape(-1, 1)
    s.solve(A, u, b)
    assert np.array_equal(np.array(u.data_ro).ravel(), u1)
    Am = A.values
    for bv, uv in ((b1, u1), (b2, u2)):
        assert np.abs(Am @ uv - bv).max() <= 1e-8 * np.abs(bv).max()


def test_oracle_pcg_matches_scipy():
    """The CG oracle (PETSc KSPCG with Jacobi, preconditioned norm) against
    scipy's preconditioned CG on an SPD system: the same iteration count and
    the same preconditioned residual norms |D^-1 (b - A x_k)| per
    iteration."""
    nx = 14
    T = sps.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(nx, nx))
    L0 = sps.kron(sps.identity(nx), T) + sps.kron(T, sps.identity(nx))
    n = nx * nx
    D = sps.diags(np.random.default_rng(9).uniform(0.5, 2.0, n))
    A = (D @ L0 @ D).tocsr()                           # SPD, Jacobi not trivial
    A.sort_indices()
    b = np.random.default_rng(10).random(n)
    x, its, rn, reason, hist = ora('pcg', A, b, rtol=1e-10)
    assert reason == 2
    dinv = 1.0 / A.diagonal()
    xs = []
    sl.cg(A, b, rtol=1e-14, atol=0.0, maxiter=its, M=sps.diags(dinv),
          callback=lambda xk: xs.append(xk.copy()))
    assert len(xs) == its
    z = np.array([np.linalg.norm(dinv * (b - A @ xk)) for xk in xs])
    assert close_history(hist[1:], z, 1e-8 * hist[0]) < 1e-6
    assert np.abs(x - np.linalg.solve(A.toarray(), b)).max() <= 1e-8 * np.abs(x).max()


@pytest.mark.gpu
def test_op2_assembled_spd_cg(hip):
    """CG on an assembled SPD operator (P1 mass + Laplace with a lognormal
    coefficient), hip against the oracle: same reason, iterations equal or
    one apart, residual histories within 1e-5, solutions within 1e-8."""
    from pyop2_amd.workloads import fem, meshes
    mesh = meshes.TriMesh(32, seed=6)
    rng = np.random.default_rng(6)
    kap, fv = rng.lognormal(0.0, 1.0, mesh.nvert), rng.random(mesh.nvert)

    def run(W):
        nodes = W.Set(mesh.nvert, 1)
        vnodes = W.Set(mesh.nvert, 2)
        el = W.Set(mesh.nelem, 1)
        en = W.Map(el, nodes, 3, mesh.elem_node)
        ev = W.Map(el, vnodes, 3, mesh.elem_node)
        A = W.Mat(W.Sparsity((en, en)), np.float64)
        x = W.Dat(vnodes, mesh.coords.copy(), np.float64)
        k = W.Dat(nodes, kap.copy(), np.float64)
        f = W.Dat(nodes, fv.copy(), np.float64)
        b = W.Dat(nodes, np.zeros(mesh.nvert), np.float64)
        u = W.Dat(nodes, np.zeros(mesh.nvert), np.float64)
        W.par_loop(W.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
                   A((en[W.i[0]], en[W.i[1]]), W.INC), x(ev, W.READ))
        W.par_loop(W.Kernel(fem.LAPLACE_TRI, 'laplace_tri'), el(3, 3),
                   A((en[W.i[0]], en[W.i[1]]), W.INC), x(ev, W.READ), k(en, W.READ))
        W.par_loop(W.Kernel(fem.RHS_TRI, 'rhs_tri'), el(3), b(en[W.i[0]], W.INC),
                   x(ev, W.READ), f(en, W.READ))
        s = W.Solver(relative_tolerance=1e-12, maximum_iterations=5000,
                     monitor_convergence=True)
        s.solve(A, u, b)
        return (np.array(u.data_ro).ravel(), s.getIterationNumber(), s.getConvergedReason(),
                np.array(s.reshist))

    ug, ig, rg, hg = run(hip)
    uo, io, ro, ho = run(seq)
    assert rg == ro and rg > 0 and abs(ig - io) <= 1
    assert close_history(hg, ho, 1e-8 * ho[0]) < 1e-5
    assert np.abs(ug - uo).max() <= 1e-8 * np.abs(uo).max()


# -- no preconditioner (PETSc's PCNONE, preconditioner='none') --------------------

@pytest.mark.parametrize('method', ['gmres', 'bcgs', 'pcg'])
def test_oracle_no_preconditioner_matches_scipy(method):
    """Without a preconditioner the residual norms are |b - A x_k|: the
    oracle's histories against scipy's unpreconditioned solvers."""
    if method == 'pcg':
        nx = 12
        T = sps.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(nx, nx))
        A = (sps.kron(sps.identity(nx), T) + sps.kron(T, sps.identity(nx))).tocsr()
    else:
        A = convdiff(10, seed=4)
    A.sort_indices()
    b = np.random.default_rng(12).random(A.shape[0])
    x, its, rn, reason, hist = ora(method, A, b, rtol=1e-10, jacobi=0)
    assert reason == 2
    got = []
    if method == 'gmres':
        sl.gmres(A, b, rtol=1e-10, atol=0.0, restart=30, maxiter=500, callback=got.append,
                 callback_type='pr_norm')
        ref = np.array(got) * np.linalg.norm(b)
    else:
        solver = sl.cg if method == 'pcg' else sl.bicgstab
        solver(A, b, rtol=1e-14, atol=0.0, maxiter=its, callback=lambda xk: got.append(xk.copy()))
        ref = np.array([np.linalg.norm(b - A @ xk) for xk in got])
    assert len(ref) == its
    assert close_history(hist[1:], ref, 1e-8 * hist[0]) < 1e-5
    assert np.abs(x - np.linalg.solve(A.toarray(), b)).max() <= 1e-8 * np.abs(x).max()


@pytest.mark.gpu
@pytest.mark.parametrize('method', METHODS)
def test_hip_no_preconditioner(method):
    A = convdiff(24, seed=24)
    b = np.random.default_rng(24).random(A.shape[0])
    assert_parity(method, A, b, rtol=1e-12, jacobi=0)


@pytest.mark.gpu
@pytest.mark.parametrize('linear_solver', ['cg', 'gmres', 'bicgstab'])
def test_op2_no_preconditioner(hip, linear_solver):
    """Solver(preconditioner='none') on the hip backend against the oracle
    on the reference's 2-triangle-free SPD case: a P1 mass matrix."""
    from pyop2_amd.workloads import fem, meshes
    mesh = meshes.TriMesh(12, seed=8)
    fv = np.random.default_rng(8).random(mesh.nvert)

    def run(W):
        nodes, vnodes = W.Set(mesh.nvert, 1), W.Set(mesh.nvert, 2)
        el = W.Set(mesh.nelem, 1)
        en = W.Map(el, nodes, 3, mesh.elem_node)
        ev = W.Map(el, vnodes, 3, mesh.elem_node)
        A = W.Mat(W.Sparsity((en, en)), np.float64)
        x = W.Dat(vnodes, mesh.coords.copy(), np.float64)
        f = W.Dat(nodes, fv.copy(), np.float64)
        b = W.Dat(nodes, np.zeros(mesh.nvert), np.float64)
        u = W.Dat(nodes, np.zeros(mesh.nvert), np.float64)
        W.par_loop(W.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
                   A((en[W.i[0]], en[W.i[1]]), W.INC), x(ev, W.READ))
        W.par_loop(W.Kernel(fem.RHS_TRI, 'rhs_tri'), el(3), b(en[W.i[0]], W.INC),
                   x(ev, W.READ), f(en, W.READ))
        s = W.Solver(linear_solver=linear_solver, preconditioner='none',
                     relative_tolerance=1e-12, maximum_iterations=5000)
        s.solve(A, u, b)
        return np.array(u.data_ro).ravel(), s.getIterationNumber(), s.getConvergedReason()

    ug, ig, rg = run(hip)
    uo, io, ro = run(seq)
    assert rg == ro and rg > 0 and abs(ig - io) <= 1
    assert np.abs(ug - fv).max() <= 1e-8 and np.abs(ug - uo).max() <= 1e-9
