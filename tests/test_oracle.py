"""Pins the CPU oracle (oracle/tsf_oracle.cpp) before it is trusted as the checker.

The reference holds no function-level golden vectors (its tests are
system-level quality bars on tests/data/synthetic/dataset_1), so every oracle
function is pinned here by what the mathematics fixes -- closed forms,
invariants, eigen-decompositions from numpy, a direct sparse solve, and two
independent implementations of the pyramid order -- chosen so that a
plausible mistake (dropped term, wrong sign, wrong index, swapped operand)
fails at least one check.

Pinning status per oracle function:
  blur_mask / blur ........ sum = 1, sigma->size table, constant and linear
                            invariance, separable scipy equivalent
  to_float ................ closed form
  vector_field ............ linear ramps (both Fmax branches), edge clamp, sign
  gvf_step / gvf_euler .... cosine eigenmode of the mirror-BC Laplacian,
                            data-term closed form, constant fixed point, mu = 0
  gvf_fmg ................. converges to the direct sparse solution of the
                            reference's discrete system; agrees with the Euler
                            steady state
  eigen3 .................. numpy eigh spectra (sorted by |lambda|),
                            orthonormality, diagonal/permuted cases
  hessian_eigen ........... linear fields (exact Jacobian), normalisation by w
  trilinear ............... exact on linear fields incl. the -0.5 shift, clamp
  tdf_circle .............. uniform-field closed form (both signs), linear
                            inward field: TDF = k*rmax*mean(c^2+s^2), radius
                            rmax; table entries equal the reference's literals
  combine / candidates .... closed form
  hp_compact .............. sort-by-key == literal pyramid descent, key digits
  candidate_nms ........... constructed cross-section / axial neighbours
  cube_max ................ independent numpy reduction
  do_eigen ................ equals hessian_eigen(normalize) with w = |F|
  sn16_store/load/round ... codes of anchor values (ties to even, saturation,
                            NaN), every one of the 65536 codes, numpy
                            restatement on random values, monotonicity
  unorm16_round ........... anchor codes, numpy restatement, order preserved
  vector_field_sn16 ....... equals the pinned float field rounded to codes
  gvf_euler_sn16 .......... equals the pinned float step with every iterate
                            rounded (and w = round(|v|)); constant fixed
                            point; bounded where the float scheme diverges
  link_centerpoints ....... known answers: a straight chain links each inner
                            vertex to its two neighbours (duplicates dropped),
                            a right angle and a zero TDF link nothing; the
                            literal 2D pyramid descent lists the edges sorted
                            by the key sum (2 x_b + y_b) 4^b
  graph_components ........ scipy connected_components (labels = smallest
                            index of the component), sizes by bincount
  rasterize_tree_edges .... known voxels of axis and diagonal edges, the tree
                            size threshold
  segment_inverse_gradient  with no growth (a uniform field) the closing of
                            the centerline equals scipy's maximum then
                            minimum filter (zero / nearest borders), two
                            passes; an inward field rising to a tube wall
                            grows the tube's cylinder and no further
  segment_spheres ......... the lattice points of the open ball, clipped
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def rng(seed=0):
    return np.random.default_rng(seed)


# ------------------------------------------------------------------ A1 blur
@pytest.mark.parametrize("sigma,m", [(0.25, 1), (0.5, 1), (0.75, 2), (1.0, 2), (1.5, 3), (2.0, 4), (2.6, 5),
                                     (10.0, 5)])
def test_blur_mask_size_and_sum(orc, sigma, m):
    mm, mask = orc.blur_mask(sigma)
    assert mm == m
    assert mask.size == (2 * m + 1) ** 3
    assert abs(float(mask.astype(np.float64).sum()) - 1.0) < 1e-5  # float accumulation of the sum
    M = 2 * m + 1
    k = mask.reshape(M, M, M)  # [c][b][a]
    assert np.allclose(k, k[::-1, :, :]) and np.allclose(k, k[:, ::-1, :]) and np.allclose(k, k[:, :, ::-1])
    # separable: w(a,b,c) proportional to g(a) g(b) g(c)
    g = np.exp(-np.arange(-m, m + 1) ** 2 / (2 * sigma * sigma))
    sep = np.einsum("i,j,k->ijk", g, g, g)
    assert np.allclose(k, sep / sep.sum(), rtol=1e-5, atol=1e-9)


def test_blur_constant_and_linear(orc):
    # the float mask sums to 1 only up to float rounding; invariance is w.r.t. that sum
    wsum = float(orc.blur_mask(1.0)[1].astype(np.float64).sum())
    vol = np.full((12, 11, 10), 3.25, np.float64)
    assert np.allclose(orc.blur(vol, 1.0, "f64"), 3.25 * wsum, rtol=1e-12)
    z, y, x = np.meshgrid(np.arange(12), np.arange(11), np.arange(10), indexing="ij")
    lin = 0.3 * x - 0.2 * y + 0.1 * z + 1.0
    out = orc.blur(lin, 1.0, "f64")  # m = 2: exact away from a 2-voxel border
    assert np.allclose(out[2:-2, 2:-2, 2:-2], wsum * lin[2:-2, 2:-2, 2:-2], atol=1e-12)


def test_blur_matches_separable_scipy(orc):
    v = rng(1).standard_normal((9, 10, 11))
    sigma = 0.9
    m, mask = orc.blur_mask(sigma)
    g = np.exp(-np.arange(-m, m + 1) ** 2 / (2 * np.float32(sigma) ** 2))
    g = g / g.sum()
    ref = v
    for ax in range(3):
        ref = ndi.correlate1d(ref, g, axis=ax, mode="nearest")
    out = orc.blur(v, sigma, "f64")
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------------ A2 toFloat
def test_to_float_closed_form(orc):
    d = np.array([-5, 0, 10, 50, 75, 100, 255], np.float64)
    out = orc.to_float(d, 0.0, 75.0, "f64")
    exp = np.clip(d, 0, 75) / 75.0
    assert np.allclose(out, exp, rtol=1e-7)


# ------------------------------------------------------------------ A3 vector field
@pytest.mark.parametrize("a,fmax,sign", [(0.05, 0.2, 1), (0.5, 0.2, 1), (0.05, 0.2, -1), (0.3, 0.1, -1)])
def test_vector_field_linear_ramp(orc, a, fmax, sign):
    Z, Y, X = 6, 7, 8
    x = np.arange(X, dtype=np.float64)
    vol = np.broadcast_to(a * x, (Z, Y, X)).copy()
    F = orc.vector_field(vol, fmax, sign, "f64")
    g = sign * a
    exp_x = g / fmax if abs(g) < fmax else np.sign(g)
    assert np.allclose(F[0][:, :, 1:-1], exp_x, rtol=1e-6)
    # clamp-to-edge: at x = 0 the difference is 0.5*(I1 - I0) = a/2
    ge = sign * 0.5 * a
    exp_e = ge / fmax if abs(ge) < fmax else np.sign(ge)
    assert np.allclose(F[0][:, :, 0], exp_e, rtol=1e-6)
    assert np.allclose(F[1], 0) and np.allclose(F[2], 0) and np.all(F[3] == 1)


def test_vector_field_direction_components(orc):
    Z, Y, X = 6, 6, 6
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    vol = 3.0 * x + 4.0 * y + 12.0 * z  # |grad| = 13 > Fmax: unit vector
    F = orc.vector_field(vol.astype(np.float64), 0.2, 1, "f64")
    c = (slice(1, -1),) * 3
    assert np.allclose(F[0][c], 3 / 13) and np.allclose(F[1][c], 4 / 13) and np.allclose(F[2][c], 12 / 13)


# ------------------------------------------------------------------ A4 GVF
def _cos_mode(shape, m):
    Z, Y, X = shape
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    v = (np.cos(np.pi * m[0] * (x - 1) / (X - 3)) * np.cos(np.pi * m[1] * (y - 1) / (Y - 3)) *
         np.cos(np.pi * m[2] * (z - 1) / (Z - 3)))
    lam = 1 + 2 * float(np.float32(0.1)) * sum(np.cos(np.pi * mk / (N - 3)) - 1 for mk, N in zip(m, (X, Y, Z)))
    return v, lam


def test_gvf_step_cosine_eigenmode(orc):
    shape = (10, 9, 12)
    v, lam = _cos_mode(shape, (2, 1, 3))
    f = np.zeros((3,) + shape)
    V = np.stack([v, 0.5 * v, -v])
    out = orc.gvf_step(f, V, 0.1, "f64")
    assert np.allclose(out, lam * V, rtol=0, atol=1e-13)
    # several steps: lam^n
    W = V
    for _ in range(5):
        W = orc.gvf_step(f, W, 0.1, "f64")
    assert np.allclose(W, lam ** 5 * V, atol=1e-12)


def test_gvf_step_data_term(orc):
    shape = (6, 7, 8)
    f = rng(2).uniform(-1, 1, (3,) + shape)
    out = orc.gvf_step(f, np.zeros_like(f), 0.1, "f64")
    # v = 0: v' = f(q) |f(q)|^2 at q = mirror(w)
    def remap(N):
        i = np.arange(N)
        i = np.where(i == 0, 2, i)
        return np.where(i >= N - 1, N - 3, i)
    qz, qy, qx = np.ix_(remap(shape[0]), remap(shape[1]), remap(shape[2]))
    fq = f[:, qz, qy, qx]
    exp = fq * (fq ** 2).sum(0)
    assert np.allclose(out, exp, atol=1e-14)


def test_gvf_fixed_points(orc):
    shape = (8, 8, 8)
    c = np.array([0.3, -0.2, 0.5])
    F = np.concatenate([np.broadcast_to(c[:, None, None, None], (3,) + shape), np.ones((1,) + shape)])
    V = orc.gvf_euler(F, 0.1, 7, "f64")
    assert np.allclose(V[:3], F[:3], atol=1e-14)
    assert np.allclose(V[3], np.linalg.norm(c))
    # mu = 0 with v0 = f: v stays f
    F2 = np.concatenate([rng(3).uniform(-1, 1, (3,) + shape), np.ones((1,) + shape)])
    V2 = orc.gvf_euler(F2, 0.0, 9, "f64")
    q = (slice(1, -1),) * 3
    assert np.allclose(V2[:3][(slice(None),) + q], F2[:3][(slice(None),) + q], atol=1e-14)


def test_gvf_finish_zero_guard(orc):
    F = np.zeros((4, 6, 6, 6))
    V = orc.gvf_euler(F, 0.1, 2, "f64")
    assert np.all(V[3] == 1.0)


# ------------------------------------------------------------------ A7 FMG
def _direct_solution(F, mu, comp):
    """Solve mu*lap(v) - s v = -s F_c on interior voxels with v(0) = v(2),
    v(N-1) = v(N-3) (the fixed point of the reference's defect correction)."""
    _, Z, Y, X = F.shape
    n = X * Y * Z
    s = (F[:3] ** 2).sum(0).ravel()
    idx = np.arange(n).reshape(Z, Y, X)

    def rm(p, N):
        return 2 if p == 0 else (N - 3 if p >= N - 1 else p)

    rows, cols, vals, b = [], [], [], np.zeros(n)
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                i = idx[z, y, x]
                qx, qy, qz = rm(x, X), rm(y, Y), rm(z, Z)
                if (qx, qy, qz) != (x, y, z):  # boundary: v(w) - v(q) = 0
                    rows += [i, i]
                    cols += [i, idx[qz, qy, qx]]
                    vals += [1.0, -1.0]
                    continue
                rows.append(i); cols.append(i); vals.append(-6 * mu - s[i])
                for dz, dy, dx in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
                    rows.append(i); cols.append(idx[z + dz, y + dy, x + dx]); vals.append(mu)
                b[i] = -s[i] * F[comp].ravel()[i]
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return spla.spsolve(A.tocsc(), b).reshape(Z, Y, X)


def test_fmg_converges_to_direct_solution(orc):
    shape = (16, 16, 16)
    F = np.zeros((4,) + shape)
    F[:3] = rng(4).uniform(-1, 1, (3,) + shape) * (rng(5).uniform(0, 1, shape) > 0.7)
    F[3] = 1
    mu = float(np.float32(0.1))  # the oracle takes mu as float
    # the defect correction converges linearly (about 0.77 per step with the
    # reference's smoother), so run it long enough to reach 1e-8
    V = orc.gvf_fmg(F, mu, 100, prec="f64")
    for c in range(3):
        ref = _direct_solution(F, mu, c)
        err = np.abs(V[c] - ref).max() / np.abs(ref).max()
        assert err < 1e-8, (c, err)
    assert np.allclose(V[3], np.sqrt((V[:3] ** 2).sum(0)))


def test_fmg_equals_euler_steady_state(orc):
    shape = (12, 12, 12)
    F = np.zeros((4,) + shape)
    F[:3] = rng(6).uniform(-1, 1, (3,) + shape) / np.sqrt(3)  # |f| <= 1 keeps the explicit scheme stable
    F[3] = 1
    Vf = orc.gvf_fmg(F, 0.1, 100, prec="f64")
    Ve = orc.gvf_euler(F, 0.1, 4000, prec="f64")
    assert np.abs(Vf[:3] - Ve[:3]).max() < 1e-8


def test_fmg_float_close_to_double(orc):
    shape = (16, 16, 16)
    F = np.zeros((4,) + shape, np.float32)
    F[:3] = rng(7).uniform(-1, 1, (3,) + shape)
    F[3] = 1
    a = orc.gvf_fmg(F, 0.15, 3, prec="f32")
    b = orc.gvf_fmg(F, 0.15, 3, prec="f64")
    assert np.linalg.norm(a[:3] - b[:3]) / np.linalg.norm(b[:3]) < 1e-5


# ------------------------------------------------------------------ A9 eigen
def _check_eigen(orc, A, prec, tol):
    d, V = orc.eigen3(A, prec)
    w, U = np.linalg.eigh(A)
    order = np.argsort(np.abs(w), kind="stable")
    assert np.allclose(d, w[order], atol=tol * max(1, np.abs(w).max()))
    assert np.all(np.diff(np.abs(d)) >= -tol)
    assert np.allclose(V.T @ V, np.eye(3), atol=tol * 10)
    assert np.allclose(A @ V, V * d[None, :], atol=tol * 10 * max(1, np.abs(w).max()))


def test_eigen_random_symmetric(orc):
    g = rng(8)
    for _ in range(200):
        B = g.standard_normal((3, 3))
        A = B + B.T
        _check_eigen(orc, A, "f64", 1e-12)
        _check_eigen(orc, A.astype(np.float32), "f32", 3e-5)


def test_eigen_known_spectra(orc):
    g = rng(9)
    for lam in ([1.0, -2.0, 3.0], [0.0, 0.0, 1.0], [-1e-3, 1e-3, 5.0], [2.0, 2.0, 2.0], [0.0, 0.0, 0.0]):
        Q, _ = np.linalg.qr(g.standard_normal((3, 3)))
        A = Q @ np.diag(lam) @ Q.T
        A = 0.5 * (A + A.T)
        d, V = orc.eigen3(A, "f64")
        assert np.allclose(np.sort(d), np.sort(lam), atol=1e-12)
        assert np.allclose(np.abs(d), np.sort(np.abs(lam)), atol=1e-12)


def test_eigen_diagonal_exact(orc):
    d, V = orc.eigen3(np.diag([3.0, -1.0, 2.0]), "f64")
    assert np.array_equal(d, [-1.0, 2.0, 3.0])
    # eigenvectors are +-unit axes: lambda -1 -> y, 2 -> z, 3 -> x
    assert np.allclose(np.abs(V), [[0, 0, 1], [1, 0, 0], [0, 1, 0]])


# ------------------------------------------------------------------ A8 Hessian
def test_hessian_linear_field(orc):
    Z, Y, X = 7, 8, 9
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    J = np.array([[0.02, 0.01, -0.03], [0.01, 0.05, 0.02], [-0.03, 0.02, -0.04]])  # symmetric
    P = np.stack([x, y, z]).astype(np.float64)
    F = np.zeros((4, Z, Y, X))
    F[:3] = np.einsum("ij,jzyx->izyx", J, P)
    F[3] = 2.0
    lam, vec = orc.hessian_eigen(F, 0, prec="f64")
    w = np.linalg.eigvalsh(J)
    exp = w[np.argsort(np.abs(w))]
    assert np.allclose(lam[:, 3, 4, 4], exp, atol=1e-12)
    lam_n, _ = orc.hessian_eigen(F, 1, prec="f64")  # every sample / w = 2
    assert np.allclose(lam_n[:, 3, 4, 4], exp / 2, atol=1e-12)


# ------------------------------------------------------------------ trilinear
def test_trilinear_linear_field_and_shift(orc):
    Z, Y, X = 6, 7, 8
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    F = np.zeros((4, Z, Y, X))
    F[0] = 1.5 * x + 0.25
    F[1] = -0.5 * y
    F[2] = 2.0 * z - 1
    for (u, v, w) in [(3.2, 2.9, 3.0), (1.5, 1.5, 1.5), (4.75, 3.1, 2.6)]:
        s = orc.trilinear(F, u, v, w)
        # the linear filter samples texel centres at i + 0.5: value at u is f(u - 0.5)
        assert np.allclose(s, [1.5 * (u - 0.5) + 0.25, -0.5 * (v - 0.5), 2.0 * (w - 0.5) - 1], atol=1e-12)
    # integer coordinate -> mean of two neighbours; far outside -> clamped edge value
    s = orc.trilinear(F, 3.0, 3.5, 2.5)
    assert np.isclose(s[0], 0.5 * (F[0, 2, 3, 2] + F[0, 2, 3, 3]))
    s = orc.trilinear(F, -5.0, 3.5, 2.5)
    assert np.isclose(s[0], F[0, 2, 3, 0])


# ------------------------------------------------------------------ A11 TDF
def test_circle_tables_match_reference_literals(orc):
    c, s = orc.circle_tables()
    # literal entries of the reference's tables (kernels_no_3d_write.cl:1500-1501)
    assert c[0] == np.float32(1.0) and c[1] == np.float32(0.540302) and c[11] == np.float32(0.0044257)
    assert c[22] == np.float32(-0.999961) and c[31] == np.float32(0.914742)
    assert s[1] == np.float32(0.841471) and s[22] == np.float32(-0.00885131) and s[11] == np.float32(-0.99999)
    assert np.allclose(c, np.cos(np.arange(32)), atol=5e-6) and np.allclose(s, np.sin(np.arange(32)), atol=5e-6)
    assert abs(float(c.astype(np.float64).sum()) - 0.5875767) < 1e-6
    assert abs(float(s.astype(np.float64).sum()) - (-0.1239863)) < 1e-6


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_tdf_uniform_field(orc, sign):
    shape = (8, 8, 8)
    cvec = sign * np.array([0.3, -0.4, 0.2])
    F = np.zeros((4,) + shape)
    F[:3] = cvec[:, None, None, None]
    F[3] = 1
    tdf, rad = orc.tdf_circle(F, 1.0, 3.0, 0.5, "f64")
    # zero Jacobian -> tred2/tql2 of the zero matrix returns the identity basis
    lam, vec = orc.hessian_eigen(F, 0, prec="f64")
    e2 = vec[3:6, 4, 4, 4]
    e3 = vec[6:9, 4, 4, 4]
    c, s = orc.circle_tables()
    S = -(cvec @ e3 * c.astype(np.float64).sum() + cvec @ e2 * s.astype(np.float64).sum()) / 32
    if S > 0:
        assert np.isclose(tdf[4, 4, 4], S, rtol=1e-12) and rad[4, 4, 4] == 1.0
    else:
        assert tdf[4, 4, 4] == 0 and rad[4, 4, 4] == 0
    assert np.all(tdf == tdf[4, 4, 4])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tdf_split_reduces_to_tdf(orc, prec):
    """Hessian and samples from one field: the reference's filter, bit for bit."""
    F = np.concatenate([rng(30).uniform(-1, 1, (3, 10, 11, 12)), rng(31).uniform(0.5, 1.5, (1, 10, 11, 12))])
    for rmin, rmax in ((1.0, 3.0), (2.5, 4.0)):
        t0, r0 = orc.tdf_circle(F, rmin, rmax, 0.5, prec)
        t1, r1 = orc.tdf_circle_split(F, F, rmin, rmax, 0.5, prec)
        assert np.array_equal(t0, t1) and np.array_equal(r0, r1)


def test_tdf_split_hessian_and_samples_from_their_fields(orc):
    """Hessian field: linear, so its Jacobian is a known non-axis-aligned
    symmetric matrix at every interior voxel; sample field: uniform c.  Then
    S(r) = -c . (e3 sum c_j + e2 sum s_j) / 32 with e2, e3 of the linear
    field's Hessian at every radius (a uniform field samples to c), so the
    response is that value at rmin, or 0; taking the Hessian from the sample
    field (zero Jacobian, identity basis) or the samples from the Hessian
    field gives other values."""
    shape = (12, 12, 12)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    A = np.array([[0.30, 0.10, -0.05], [0.10, -0.20, 0.15], [-0.05, 0.15, 0.05]])  # symmetric Jacobian
    FH = np.zeros((4,) + shape)
    P = np.stack([x, y, z])  # d F_i / d x_j = A[i, j]
    for i in range(3):
        FH[i] = sum(A[i, j] * P[j] for j in range(3))
    FH[3] = 1
    c, s = orc.circle_tables()
    for cvec in (np.array([0.3, -0.4, 0.2]), np.array([-0.3, 0.4, -0.2])):
        FS = np.zeros((4,) + shape)
        FS[:3] = cvec[:, None, None, None]
        FS[3] = 1
        tdf, rad = orc.tdf_circle_split(FH, FS, 1.0, 3.0, 0.5, "f64")
        lam, vec = orc.hessian_eigen(FH, 0, prec="f64")
        e2, e3 = vec[3:6, 6, 6, 6], vec[6:9, 6, 6, 6]
        assert np.allclose(np.abs(np.sort(np.abs(lam[:, 6, 6, 6]))), np.sort(np.abs(np.linalg.eigvalsh(A))))
        S = -(cvec @ e3 * c.astype(np.float64).sum() + cvec @ e2 * s.astype(np.float64).sum()) / 32
        if S > 0:
            assert np.isclose(tdf[6, 6, 6], S, rtol=1e-12) and rad[6, 6, 6] == 1.0
        else:
            assert tdf[6, 6, 6] == 0 and rad[6, 6, 6] == 0
        ident = -(cvec[2] * c.astype(np.float64).sum() + cvec[1] * s.astype(np.float64).sum()) / 32
        assert not np.isclose(max(S, 0.0), max(ident, 0.0)) or S <= 0 and ident <= 0


@pytest.mark.parametrize("prec,rtol", [("f64", 1e-12), ("f32", 2e-6)])
def test_tdf_mixed_known_answer(orc, prec, rtol):
    """The precision form's fused sampling on the linear-Hessian / uniform
    sample field pair: a uniform field samples to c exactly (the weights sum
    to one in any rounding order only to within an ulp, hence rtol), so the
    response is S(rmin) = -c . (e3 sum c_j + e2 sum s_j) / 32 or 0.  The sum
    is the same at every radius, so in float a rounding-level rise may carry
    the search past rmin: the f32 radius is any on the grid."""
    shape = (12, 12, 12)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    A = np.array([[0.30, 0.10, -0.05], [0.10, -0.20, 0.15], [-0.05, 0.15, 0.05]])
    FH = np.zeros((4,) + shape)
    P = np.stack([x, y, z])
    for i in range(3):
        FH[i] = sum(A[i, j] * P[j] for j in range(3))
    FH[3] = 1
    c, s = orc.circle_tables()
    lam, vec = orc.hessian_eigen(FH, 0, prec="f64")
    e2, e3 = vec[3:6, 6, 6, 6], vec[6:9, 6, 6, 6]
    for cvec in (np.array([0.3, -0.4, 0.2]), np.array([-0.3, 0.4, -0.2])):
        FS = np.zeros((4,) + shape)
        FS[:3] = cvec[:, None, None, None]
        FS[3] = 1
        tdf, rad = orc.tdf_circle_mixed(FH, FS, 1.0, 3.0, 0.5, prec)
        S = -(cvec @ e3 * c.astype(np.float64).sum() + cvec @ e2 * s.astype(np.float64).sum()) / 32
        if S > 0:
            assert np.isclose(tdf[6, 6, 6], S, rtol=rtol)
            assert rad[6, 6, 6] == 1.0 if prec == "f64" else rad[6, 6, 6] in (1.0, 1.5, 2.0, 2.5, 3.0)
        else:
            assert tdf[6, 6, 6] == 0 and rad[6, 6, 6] == 0


@pytest.mark.parametrize("prec,rtol", [("f64", 1e-11), ("f32", 1e-4)])
def test_tdf_mixed_linear_field_known_answer(orc, prec, rtol):
    """The fused sampling on a linear sample field FS(u) = B u + c0, which
    trilinear interpolation reproduces whatever its weights (so a wrong
    weight, a swapped plane or a dropped tap shows): the sample at
    u = p - 0.5 + r V_j is B u + c0, the response at radius r is
    S(r) = -(1/32) sum_j V_j . (B (p - 0.5 + r V_j) + c0), and the search
    keeps the first maximum over r = 1, 1.5, .., 3 (0 if S(1) <= 0).  The
    cases cover a response climbing to rmax, one falling from rmin and none."""
    shape = (12, 12, 12)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    A = np.array([[0.30, 0.10, -0.05], [0.10, -0.20, 0.15], [-0.05, 0.15, 0.05]])
    FH = np.zeros((4,) + shape)
    P = np.stack([x, y, z])
    for i in range(3):
        FH[i] = sum(A[i, j] * P[j] for j in range(3))
    FH[3] = 1
    c, s = orc.circle_tables()
    lam, vec = orc.hessian_eigen(FH, 0, prec="f64")
    e2, e3 = vec[3:6, 6, 6, 6], vec[6:9, 6, 6, 6]
    V = np.outer(c.astype(np.float64), e3) + np.outer(s.astype(np.float64), e2)
    q = np.full(3, 6.0) - 0.5
    B0 = np.array([[-0.12, 0.03, 0.01], [0.02, -0.07, -0.04], [0.05, 0.01, -0.10]])
    seen = set()
    for B, c0 in ((B0, np.array([0.2, -0.1, 0.05])), (-B0, np.array([6.0, 0.0, 0.0])),
                  (-B0, np.array([1.5, -0.9, -0.7]))):
        FS = np.zeros((4,) + shape)
        for i in range(3):
            FS[i] = sum(B[i, j] * P[j] for j in range(3)) + c0[i]
        FS[3] = 1
        tdf, rad = orc.tdf_circle_mixed(FH, FS, 1.0, 3.0, 0.5, prec)
        best, rbest = 0.0, 0.0
        for r in np.arange(1.0, 3.01, 0.5):
            S = -np.mean(np.einsum("ji,ji->j", V, (q + r * V) @ B.T + c0))
            if S > best:
                best, rbest = S, r
            else:
                break
        seen.add(rbest)
        assert rad[6, 6, 6] == rbest
        assert np.isclose(tdf[6, 6, 6], best, rtol=rtol, atol=0)
    assert seen == {0.0, 1.0, 3.0}


def test_tdf_mixed_fused_is_as_close_to_fp64(orc):
    """On a GVF field (FH = FS, no 16-bit rounding): the fused float
    sampling agrees with the FP64 filter to float rounding on the voxels
    whose radius matches, at least as closely as the reference-order float
    filter, and changes no more radii; and the two float forms differ (the
    fused arithmetic is in use)."""
    from tsf import synth
    vol, _ = synth.tube_volume((24, 28, 32), 5, 1.0, 3.0, 60, noise=0.05)
    F = orc.vector_field(orc.blur(vol, 0.5, "f64"), 0.2, 1, "f64")
    V = orc.gvf_euler(F, 0.05, 40, "f64")
    for rmin, rmax in ((0.5, 3.0), (2.5, 4.0)):
        t64, r64 = orc.tdf_circle(V, rmin, rmax, 0.5, "f64")
        tf, rf = orc.tdf_circle_mixed(V.astype(np.float32), V.astype(np.float32), rmin, rmax, 0.5, "f32")
        tu, ru = orc.tdf_circle(V.astype(np.float32), rmin, rmax, 0.5, "f32")
        resp = t64 > 0.05
        for t, r in ((tf, rf), (tu, ru)):
            same = resp & (r == r64)
            assert np.abs(t[same] - t64[same]).max() < 1e-5 * np.abs(t64[same]).max()
        same_f, same_u = resp & (rf == r64), resp & (ru == r64)
        ef = np.linalg.norm(tf[same_f] - t64[same_f]) / np.linalg.norm(t64[same_f])
        eu = np.linalg.norm(tu[same_u] - t64[same_u]) / np.linalg.norm(t64[same_u])
        print(f"r {rmin}-{rmax}: fused {ef:.2e}, reference order {eu:.2e}")
        assert ef <= 1.5 * eu and (resp & (rf != r64)).sum() <= (resp & (ru != r64)).sum() + 2
        assert not np.array_equal(tf, tu)


def test_eigen_fp64_fallback_resolves_near_degenerate_pairs(orc):
    """Symmetric matrices with two eigenvalues 1e-5 apart (relative): the
    float QL usually resolves the pair, but in its worst cases leaves the
    eigenvectors wrong by up to eps / gap; the fallback (FP64 from the same
    float matrix, rounded) matches numpy's FP64 eigh to float rounding in
    every case.  Well-separated matrices take the float result unchanged."""
    g = rng(51)
    err_f, err_fb = [], []
    for _ in range(200):
        Q, _ = np.linalg.qr(g.standard_normal((3, 3)))
        lam = np.array([0.3, 1.0, 1.0 + 1e-5]) * g.choice([-1.0, 1.0], 3)
        A = ((Q * lam) @ Q.T).astype(np.float32)
        A = (A + A.T) / 2
        w, U = np.linalg.eigh(A.astype(np.float64))
        o = np.argsort(np.abs(w))
        w, U = w[o], U[:, o]
        for fn, errs in ((lambda M: orc.eigen3(M, "f32"), err_f), (orc.eigen3_fp64_fallback, err_fb)):
            d, V = fn(A)
            errs.append(max(min(np.abs(V[:, k] - U[:, k]).max(), np.abs(V[:, k] + U[:, k]).max()) for k in range(3)))
            assert np.allclose(d, w, rtol=1e-5, atol=1e-6)
    print(f"eigenvector error: float {np.median(err_f):.1e} (max {max(err_f):.1e}), fallback {max(err_fb):.1e}")
    assert max(err_fb) < 2e-6 and max(err_f) > 100 * max(err_fb)
    for _ in range(50):
        A = g.standard_normal((3, 3)).astype(np.float32)
        A = (A + A.T) / 2
        d, V = orc.eigen3(A, "f32")
        w = np.sort(np.abs(d))
        if (min(w[1] - w[0], w[2] - w[1]) < 2 ** -9 * w[2]):
            continue
        d2, V2 = orc.eigen3_fp64_fallback(A)
        assert np.array_equal(d, d2) and np.array_equal(V, V2)


def test_tdf_linear_inward_field(orc):
    Z, Y, X = 9, 24, 24
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    k = 0.02
    c0 = 11.5  # the sampled field is f(p - 0.5): centre it between voxels 11 and 12 -> at voxel 12 the
    F = np.zeros((4, Z, Y, X))  # sampled inward field is -k (p - 12)
    F[0] = -k * (x - c0)
    F[1] = -k * (y - c0)
    F[3] = 1
    tdf, rad = orc.tdf_circle(F, 1.0, 5.0, 1.0, "f64")
    c, s = orc.circle_tables()
    m = np.mean(c.astype(np.float64) ** 2 + s.astype(np.float64) ** 2)
    # Jacobian diag(-k, -k, 0): e1 = z, e2/e3 span xy; S(r) = k r m grows with r -> radius rmax
    assert rad[4, 12, 12] == 5.0
    assert np.isclose(tdf[4, 12, 12], k * 5.0 * m, rtol=1e-9)


# ------------------------------------------------------------------ A12, A13
def test_combine_and_candidates(orc):
    g = rng(10)
    ts, rs, tl, rl = (g.uniform(0, 1, 100) for _ in range(4))
    a, b = orc.combine(ts, rs, tl, rl, "f64")
    take = tl < ts
    assert np.array_equal(a, np.where(take, ts, tl)) and np.array_equal(b, np.where(take, rs, rl))
    m = orc.candidates(ts, 0.5, "f64")
    assert np.array_equal(m, (ts >= 0.5).astype(np.uint8))


# ------------------------------------------------------------------ A14, A15
def test_hp_key_digits(orc):
    assert orc.hp_key(1, 0, 0) == 1 and orc.hp_key(0, 0, 1) == 2 and orc.hp_key(0, 1, 0) == 4
    assert orc.hp_key(2, 0, 0) == 8 and orc.hp_key(0, 2, 0) == 32 and orc.hp_key(1, 1, 1) == 7


@pytest.mark.parametrize("shape,p", [((16, 16, 16), 0.1), ((20, 12, 28), 0.3), ((5, 9, 7), 0.5), ((40, 36, 32), 0.02),
                                     ((8, 8, 8), 1.0), ((8, 8, 8), 0.0)])
def test_hp_sort_equals_pyramid(orc, shape, p):
    mask = (rng(11).uniform(0, 1, shape) < p).astype(np.uint8)
    a = orc.hp_compact(mask, "sort")
    b = orc.hp_compact(mask, "pyramid")
    assert a.shape[0] == mask.sum() == b.shape[0]
    assert np.array_equal(a, b)
    if a.shape[0]:
        keys = [orc.hp_key(*t) for t in a]
        assert all(k1 < k2 for k1, k2 in zip(keys, keys[1:]))
        assert np.all(mask[a[:, 2], a[:, 1], a[:, 0]] == 1)


def test_hp_single_voxel(orc):
    mask = np.zeros((12, 10, 14), np.uint8)
    mask[7, 3, 11] = 1
    assert orc.hp_compact(mask, "pyramid").tolist() == [[11, 3, 7]]


# ------------------------------------------------------------------ A16 NMS
def test_candidate_nms_cross_section(orc):
    Z, Y, X = 15, 15, 15
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    F = np.zeros((4, Z, Y, X))
    F[0] = -0.05 * (x - 6.5)
    F[1] = -0.05 * (y - 6.5)
    F[3] = 1.0  # |v| plane (the NMS reads it)
    radius = np.full((Z, Y, X), 2.0)
    tdf = np.ones((Z, Y, X))
    p = np.array([[7, 7, 7]], np.int32)
    assert orc.candidate_nms(tdf, radius, F, p, 5, "f64")[7, 7, 7] == 1
    G = F.copy()
    G[3, 7, 7, 8] = 0.5  # in the cross-section (x neighbour), smaller |v| -> reject
    assert orc.candidate_nms(tdf, radius, G, p, 5, "f64")[7, 7, 7] == 0
    H = F.copy()
    H[3, 8, 7, 7] = 0.5  # along the tube axis e1 = z -> theta = pi/2 -> keep
    assert orc.candidate_nms(tdf, radius, H, p, 5, "f64")[7, 7, 7] == 1
    K = F.copy()
    K[3, 7, 7, 9] = 0.5  # |r| = 2 = maxD -> outside the open disk -> keep
    assert orc.candidate_nms(tdf, radius, K, p, 5, "f64")[7, 7, 7] == 1


# ------------------------------------------------------------------ A17 cube max
def test_cube_max_numpy(orc):
    shape = (10, 11, 13)
    g = rng(12)
    tdf = g.uniform(0, 1, shape)
    tdf[g.uniform(0, 1, shape) < 0.05] = 0.0
    cand = (g.uniform(0, 1, shape) < 0.3).astype(np.uint8)
    cube = 3
    out = orc.cube_max(tdf, cand, cube, "f64")
    exp = np.zeros(shape, np.uint8)
    Z, Y, X = shape
    for gz in range(0, Z, cube):
        for gy in range(0, Y, cube):
            for gx in range(0, X, cube):
                blk = np.where(cand[gz:gz + cube, gy:gy + cube, gx:gx + cube] == 1,
                               tdf[gz:gz + cube, gy:gy + cube, gx:gx + cube], -1.0)
                # loop order x outer, z inner: first maximum in (x, y, z) lexicographic order
                t = np.transpose(blk, (2, 1, 0)).ravel()
                if t.max() > 0:
                    k = int(np.argmax(t))
                    bx, by, bz = np.unravel_index(k, np.transpose(blk, (2, 1, 0)).shape)
                    exp[gz + bz, gy + by, gx + bx] = 1
    assert np.array_equal(out, exp)


# ------------------------------------------------------------------ A10 doEigen
def test_do_eigen_equals_normalised_hessian(orc):
    shape = (8, 9, 10)
    F = np.zeros((4,) + shape)
    F[:3] = rng(13).uniform(-1, 1, (3,) + shape)
    F[3] = np.sqrt((F[:3] ** 2).sum(0))
    lam, vec = orc.hessian_eigen(F, 1, prec="f64")
    for p in [(3, 4, 5), (1, 1, 1), (8, 7, 6)]:
        l, e1, e2, e3 = orc.do_eigen(F[0], F[1], F[2], p, "f64")
        x, y, z = p
        assert np.allclose(l, lam[:, z, y, x], atol=1e-12)
        assert np.allclose(e1, vec[0:3, z, y, x], atol=1e-12)
        assert np.allclose(e3, vec[6:9, z, y, x], atol=1e-12)


# ------------------------------------------------------------------ 16-bit formats
# The reference's 16-bit mode (kernels_no_3d_write.cl:9-21): vectors stored as
# convert_short_sat_rte(v * 32767) and read as max(-1, s / 32767); TDF values
# stored as convert_ushort_sat_rte(t * 65535) and read as t / 65535.
def _np_sn16(a):
    x = np.asarray(a, np.float32) * np.float32(32767)
    s = np.where(np.isnan(x), 0, np.rint(np.clip(np.nan_to_num(x), -32768, 32767))).astype(np.int16)
    return s


def test_sn16_codes_of_anchor_values(orc):
    v = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, np.inf, -np.inf, np.nan, 1.0 / 32767, 1e-30],
                 np.float32)
    exp = np.array([0, 0, 32767, -32767, 16384, -16384, 32767, -32768, 32767, -32768, 0, 1, 0], np.int16)
    # 0.5 * 32767 = 16383.5 is a tie: round half to even gives 16384
    assert np.array_equal(orc.sn16_store(v), exp)
    # the most negative code reads back as -1 (the max(-1, .) of the format)
    assert orc.sn16_load(np.array([-32768, -32767, 32767, 0], np.int16)).tolist() == [-1.0, -1.0, 1.0, 0.0]


def test_sn16_ties_round_to_even(orc):
    k = np.arange(-32767, 32767)
    v = ((k + 0.5) / 32767).astype(np.float32)
    tie = v * np.float32(32767) == (k + 0.5).astype(np.float32)
    assert tie.sum() > 1000
    s = orc.sn16_store(v[tie]).astype(np.int64)
    assert np.all(s % 2 == 0) and np.all(np.abs(s - (k[tie] + 0.5)) == 0.5)


def test_sn16_all_codes_and_idempotence(orc):
    codes = np.arange(-32768, 32768).astype(np.int16)
    vals = orc.sn16_load(codes)
    # float32 division is correctly rounded in numpy
    assert np.array_equal(vals, np.maximum(np.float32(-1), codes.astype(np.float32) / np.float32(32767)))
    # every code survives a store of its value (except -32768, which reads as -1)
    back = orc.sn16_store(vals)
    assert np.array_equal(back[1:], codes[1:]) and back[0] == -32767
    assert np.array_equal(orc.sn16_round(vals), vals)
    x = rng(20).uniform(-1.5, 1.5, 100000).astype(np.float32)
    assert np.array_equal(orc.sn16_store(x), _np_sn16(x))
    assert np.all(np.diff(orc.sn16_round(np.sort(x))) >= 0)


def test_unorm16_anchor_values(orc):
    v = np.array([0.0, 1.0, 0.5, -0.25, 1.5, np.nan, np.inf, 1.0 / 65535, 0.49 / 65535], np.float32)
    exp = np.array([0, 65535, 32768, 0, 65535, 0, 65535, 1, 0], np.float64)
    assert np.array_equal(orc.unorm16_round(v), (exp.astype(np.float32) / np.float32(65535)))
    x = rng(21).uniform(-0.2, 1.2, 100000).astype(np.float32)
    c = np.rint(np.clip(x * np.float32(65535), 0, 65535)).astype(np.float32)
    assert np.array_equal(orc.unorm16_round(x), c / np.float32(65535))
    # combine and candidates see the same order on rounded values as on codes
    r = orc.unorm16_round(x)
    assert np.array_equal(np.argsort(r, kind="stable"), np.argsort(c, kind="stable"))


def test_vector_field_sn16_is_rounded_float_field(orc):
    vol = rng(22).uniform(0, 1, (10, 11, 12)).astype(np.float32)
    F = orc.vector_field_sn16(vol, 0.2, 1)
    Ff = orc.vector_field(vol, 0.2, 1, "f32")
    assert np.array_equal(orc.sn16_store(F), _np_sn16(Ff))
    assert np.all(F[3] == 1.0)


def test_gvf_euler_sn16_rounds_every_iterate(orc):
    shape = (9, 10, 11)
    F = orc.sn16_round(rng(23).uniform(-0.8, 0.8, (4,) + shape))
    F[3] = 1
    mu = np.float32(0.05)
    V = F[:3].copy()
    for it in range(1, 6):
        V = orc.sn16_round(orc.gvf_step(F[:3], V, mu, "f32"))
        got = orc.gvf_euler_sn16(F, mu, it)
        assert np.array_equal(got[:3], V)
        # w = |v| in float, stored in the format (saturates at 1, 1 where 0)
        m = np.sqrt((V.astype(np.float32) ** 2).sum(0, dtype=np.float32))
        assert np.array_equal(got[3], orc.sn16_round(np.where(m > 0, m, 1).astype(np.float32)))
    # without the per-iterate rounding the iterates differ
    assert not np.array_equal(orc.gvf_euler(F, mu, 5, "f32")[:3], V)


def test_gvf_euler_sn16_fixed_point_and_saturation(orc):
    shape = (8, 8, 8)
    c = orc.sn16_round(np.array([0.3, -0.2, 0.9], np.float32))
    F = np.concatenate([np.broadcast_to(c[:, None, None, None], (3,) + shape), np.ones((1,) + shape)]).astype(
        np.float32)
    V = orc.gvf_euler_sn16(F, 0.1, 7)
    assert np.array_equal(V[:3], F[:3])
    m = np.sqrt((c.astype(np.float32) ** 2).sum(dtype=np.float32))
    assert np.all(V[3] == orc.sn16_round(np.array([m], np.float32))[0])
    # mu = 0.1 with |f| = 1 is outside the explicit scheme's stability range:
    # float iterates grow without bound, the 16-bit iterates stay within [-1, 1]
    g = rng(24)
    d = g.normal(size=(3,) + shape)
    F2 = np.concatenate([d / np.linalg.norm(d, axis=0), np.ones((1,) + shape)]).astype(np.float32)
    F2 = orc.sn16_round(F2)
    f32 = orc.gvf_euler(F2, 0.1, 300, "f32")
    s16 = orc.gvf_euler_sn16(F2, 0.1, 300)
    assert np.abs(f32[:3]).max() > 10
    assert np.abs(s16[:3]).max() <= 1.0 and np.all(np.isfinite(s16))


# ---------------------------------------------------------------- centerline graph
def _key2d(x, y):
    k = 0
    for b in range(16):
        k |= (((x >> b) & 1) * 2 + ((y >> b) & 1)) << (2 * b)
    return k


def test_link_straight_chain_known_answer(orc):
    # five vertices 3 voxels apart on a line in a uniform TDF: each inner
    # vertex links its two neighbours (180 degrees), the ends link nothing
    # (all their candidates lie in one direction); (x, y) with x > y is
    # dropped when (y, x) exists
    T = np.ones((12, 12, 20), np.float32)
    v = np.array([[2 + 3 * i, 5, 6] for i in range(5)], np.int32)
    e = orc.link_centerpoints(v, T, 25.0, 0.5)
    got = {tuple(map(int, r)) for r in e}
    assert got == {(1, 0), (1, 2), (2, 3), (3, 4)}
    keys = [_key2d(int(x), int(y)) for x, y in e]
    assert keys == sorted(keys)
    # a TDF below the minimum mean links nothing; so does a right angle
    assert len(orc.link_centerpoints(v, np.zeros_like(T), 25.0, 0.5)) == 0
    L = np.array([[5, 5, 5], [9, 5, 5], [5, 9, 5]], np.int32)
    assert len(orc.link_centerpoints(L, T, 25.0, 0.5)) == 0


def test_link_pyramid_order_is_key_order(orc):
    # many vertices on a few random polylines: the literal pyramid descent
    # lists the edges in increasing key, with no duplicate pair
    g = np.random.default_rng(11)
    pts = []
    for _ in range(6):
        p = g.uniform(8, 40, 3)
        d = g.normal(size=3)
        d /= np.linalg.norm(d)
        for k in range(12):
            p = p + 2.5 * d
            d = d + 0.2 * g.normal(size=3)
            d /= np.linalg.norm(d)
            pts.append(np.clip(np.round(p), 0, 47))
    v = np.unique(np.array(pts, np.int32), axis=0)
    T = np.ones((48, 48, 48), np.float32)
    e = orc.link_centerpoints(v, T, 25.0, 0.5)
    assert len(e) > 20
    keys = [_key2d(int(x), int(y)) for x, y in e]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    pairs = {tuple(sorted(map(int, r))) for r in e}
    assert len(pairs) == len(e)


def test_graph_components_vs_scipy(orc):
    from scipy.sparse.csgraph import connected_components
    g = np.random.default_rng(12)
    n = 300
    e = g.integers(0, n, (220, 2)).astype(np.int32)
    C, S = orc.graph_components(e, n)
    A = sp.coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    _, lab = connected_components(A, directed=False)
    mins = {}
    for i, l in enumerate(lab):
        mins.setdefault(l, i)
    assert np.array_equal(C, np.array([mins[l] for l in lab]))
    assert np.array_equal(S, np.bincount(C, minlength=n))


def test_rasterize_tree_edges_known_voxels(orc):
    v = np.array([[1, 2, 3], [5, 2, 3], [1, 6, 7]], np.int32)
    e = np.array([[0, 1], [0, 2]], np.int32)
    C = np.zeros(3, np.int32)
    S = np.array([3, 0, 0], np.int32)
    out = orc.rasterize_tree_edges(e, v, C, S, 3, (10, 10, 10))
    # axis edge: x = 1..4 (the end voxel is not set); diagonal (0, 4, 4):
    # l = round(5.657) = 6, round(4 i / 6) = 0, 1, 1, 2, 3, 3
    exp = np.zeros((10, 10, 10), np.uint8)
    exp[3, 2, 1:5] = 1
    for i in range(6):
        q = int(np.floor(4 * i / 6 + 0.5))
        exp[3 + q, 2 + q, 1] = 1
    assert np.array_equal(out, exp)
    assert not orc.rasterize_tree_edges(e, v, C, S, 4, (10, 10, 10)).any()


# ---------------------------------------------------------------- segmentation
def test_segment_closing_of_centerline_without_growth(orc):
    # w equal everywhere: no neighbour has a larger magnitude, so nothing is
    # accepted and the result is the 3x3x3 closing of the centerline
    g = np.random.default_rng(13)
    shape = (14, 16, 18)
    cl = (g.random(shape) < 0.04).astype(np.uint8)
    F = np.zeros((4,) + shape, np.float32)
    F[3] = 1.0
    seg, passes = orc.segment_inverse_gradient(cl, np.full(shape, 3.0, np.float32), F)
    exp = ndi.minimum_filter(ndi.maximum_filter(cl, size=3, mode="constant", cval=0), size=3, mode="nearest")
    assert passes == 2
    assert np.array_equal(seg, exp)


def test_segment_tube_grows_to_its_wall(orc):
    # a tube along x: the field points at the axis with magnitude rising to
    # rho = 6 and falling beyond; seeds of half size 1 on the axis grow out
    # over several passes while the magnitude rises, and never past it
    shape, P = (28, 28, 32), 6.0
    z, y, x = np.mgrid[:shape[0], :shape[1], :shape[2]].astype(np.float32)
    dy, dz = y - 13.5, z - 13.5
    rho = np.sqrt(dy * dy + dz * dz)
    m = np.where(rho <= P, rho, np.maximum(0.0, 2 * P - rho))
    F = np.zeros((4,) + shape, np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        F[1] = np.where(rho > 0, -m * dy / rho, 0.0)
        F[2] = np.where(rho > 0, -m * dz / rho, 0.0)
    w = np.sqrt(F[1] ** 2 + F[2] ** 2)
    F[3] = np.where(w > 0, w, 1.0)
    cl = np.zeros(shape, np.uint8)
    cl[13:15, 13:15, 2:30] = 1
    seg, passes = orc.segment_inverse_gradient(cl, np.full(shape, 2.0, np.float32), F)
    core = (rho <= 4.0) & (x >= 3) & (x < 29)
    cover = (seg.astype(bool) & core).sum() / core.sum()
    print(f"tube: {seg.sum()} voxels in {passes} passes, core covered {cover:.3f}")
    assert passes >= 4 and cover > 0.95
    assert not (seg.astype(bool) & (rho > P + 1.0)).any()


def test_segment_spheres_ball(orc):
    shape = (11, 12, 13)
    cl = np.zeros(shape, np.uint8)
    cl[5, 6, 6] = 1
    cl[0, 0, 0] = 1
    r = np.full(shape, 2.5, np.float32)
    seg = orc.segment_spheres(cl, r)
    z, y, x = np.mgrid[:shape[0], :shape[1], :shape[2]]
    ball = lambda cz, cy, cx: (z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2 < 6.25
    assert np.array_equal(seg.astype(bool), ball(5, 6, 6) | ball(0, 0, 0))


# ----------------------------------------------------------------- ingest
def test_volume_limits_known_answers(orc):
    """getMinimum / getMaximum: numpy's min/max on integer volumes; the
    initial values numeric_limits<T>::max() / ::min() (FLT_MIN for float);
    NaN never wins; of +0 and -0 the first stays."""
    rng = np.random.default_rng(5)
    for dt in (np.uint8, np.int8, np.uint16, np.int16):
        info = np.iinfo(dt)
        v = rng.integers(info.min, info.max, size=(5, 6, 7), endpoint=True).astype(dt)
        assert orc.volume_limits(v) == (float(v.min()), float(v.max()))
        assert orc.volume_limits(np.full((2, 2, 2), info.max, dt)) == (float(info.max), float(info.max))
    flt_min, flt_max = float(np.finfo(np.float32).tiny), float(np.finfo(np.float32).max)
    neg = -rng.random((3, 4, 5)).astype(np.float32) - 1.0
    assert orc.volume_limits(neg) == (float(neg.min()), flt_min)
    nan = np.full((2, 2, 2), np.nan, np.float32)
    assert orc.volume_limits(nan) == (flt_max, flt_min)
    nan[0, 0, 1] = 3.0
    assert orc.volume_limits(nan) == (3.0, 3.0)
    z = np.array([-0.0, 0.0, 1.0], np.float32).reshape(1, 1, 3)
    assert np.signbit(orc.volume_limits(z)[0])
    assert not np.signbit(orc.volume_limits(z[..., [1, 0, 2]])[0])


def test_crop_scan_lines_threshold_box(orc):
    """A bright box: lines_x[x] = its y extent for x inside, lines_y[y] = its x
    extent, lines_z[z] = its y extent; nothing outside; strict '>'."""
    v = np.zeros((10, 12, 14), np.uint8)
    v[2:7, 3:9, 4:11] = 200
    lx, ly, lz = orc.crop_scan_lines(v, 0, 100.0)
    ex = np.zeros(14, np.int32); ex[4:11] = 6
    ey = np.zeros(12, np.int32); ey[3:9] = 7
    ez = np.zeros(10, np.int32); ez[2:7] = 6
    assert (lx == ex).all() and (ly == ey).all() and (lz == ez).all()
    assert all((a == 0).all() for a in orc.crop_scan_lines(v, 0, 200.0))


def _lung_pattern(runs, white, black):
    return np.concatenate([np.full(n, white if w else black) for w, n in runs])


def test_crop_scan_lines_lung_automaton(orc):
    """cropDatasetLung's run-length test: white 31, black 31, white 31 counts
    (two white areas, one black); a black run of 30 counts no area; five runs
    of 31 count (more than one of each); an all-white line does not.  Lines
    along z: X, Y < 31 keep every x-line empty."""
    W, B = 0, -1000
    cases = {
        (0, 0): ([(1, 31), (0, 31), (1, 31)], True),
        (1, 0): ([(1, 31), (0, 30), (1, 32)], False),
        (2, 1): ([(1, 31), (0, 31), (1, 31), (0, 31), (1, 31)], True),
        (3, 2): ([(1, 155)], False),
    }
    Z = 155
    v = np.full((Z, 4, 5), B, np.int16)
    for (x, y), (runs, _) in cases.items():
        p = _lung_pattern(runs, W, B)
        v[:len(p), y, x] = p
    lx, ly, lz = orc.crop_scan_lines(v, 1)
    ex, ey = np.zeros(5, np.int32), np.zeros(4, np.int32)
    for (x, y), (_, inside) in cases.items():
        ex[x] += inside
        ey[y] += inside
    assert (lx == ex).all() and (ly == ey).all() and (lz == 0).all()
    # the same lines along x (Z, Y < 31): only lines_z counts
    t = np.ascontiguousarray(np.transpose(v, (2, 1, 0)))   # (x, y, z) -> shape (5, 4, 155) = (Z', Y', X')
    lx2, ly2, lz2 = orc.crop_scan_lines(t, 1)
    assert (lx2 == 0).all() and (ly2 == 0).all()
    ez = np.zeros(5, np.int32)
    for (x, y), (_, inside) in cases.items():
        ez[x] += inside
    assert (lz2 == ez).all()


def test_crop_scan_lines_lung_unsigned_limit_and_short_wrap(orc):
    """Unsigned volumes shift the limit by 1024 (white > 874); a value read
    into the kernel's short wraps (40000 -> -25536: black)."""
    runs = [(1, 31), (0, 31), (1, 31)]
    for white, inside in ((1024, 1), (875, 1), (874, 0), (30000, 1), (40000, 0)):
        v = np.zeros((93, 1, 1), np.uint16)
        v[:, 0, 0] = _lung_pattern(runs, white, 0)
        assert orc.crop_scan_lines(v, 1)[0][0] == inside, white


def test_crop_region_known_answers(orc):
    """The box: x [4, 11) grows high then low to [3, 11); y [3, 8] (the last
    qualifying slice is the exclusive end) grows to [2, 10); z [2, 6) stays.
    Middle start: z from the first slices at or below the minimum on either
    side of Z/2.  All slices qualifying: the top slice is excluded, then the
    extent grows back to the volume and is trimmed to a multiple of 4.  No
    qualifying z slice with the end start: an empty extent."""
    v = np.zeros((10, 12, 14), np.uint8)
    v[2:7, 3:9, 4:11] = 200
    assert orc.crop_region(*orc.crop_scan_lines(v, 0, 100.0), 0, False) == ((3, 2, 2), (8, 8, 4))
    lz = np.array([0] * 3 + [300] * 14 + [0] * 3, np.int32)
    full = np.full(16, 300, np.int32)
    assert orc.crop_region(full, full, lz, 200, True) == ((0, 0, 2), (16, 16, 16))
    ten = np.full(10, 300, np.int32)
    assert orc.crop_region(ten, full, full, 200, False) == ((0, 0, 0), (8, 16, 16))
    assert orc.crop_region(full, full, np.zeros(16, np.int32), 200, False) is None


# ----------------------------------------------------------------- spline TDF
def _spline_tube(shape, peak, axis, sign=-1.0):
    """A tube along `axis` (0 = z, 1 = y, 2 = x) centred in the volume: the
    field points inward (sign -1) with magnitude rho up to `peak`, then falls."""
    Z, Y, X = shape
    zz, yy, xx = np.mgrid[:Z, :Y, :X].astype(np.float64)
    dd = [zz - (Z - 1) / 2, yy - (Y - 1) / 2, xx - (X - 1) / 2]
    a, b = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
    rho = np.sqrt(dd[a] ** 2 + dd[b] ** 2)
    m = np.where(rho <= peak, rho, np.maximum(0.0, 2 * peak - rho))
    F = np.zeros((4,) + shape)
    plane = {0: 2, 1: 1, 2: 0}
    with np.errstate(invalid="ignore", divide="ignore"):
        F[plane[a]] = np.where(rho > 0, sign * m * dd[a] / rho, 0)
        F[plane[b]] = np.where(rho > 0, sign * m * dd[b] / rho, 0)
    w = np.sqrt(F[0] ** 2 + F[1] ** 2 + F[2] ** 2)
    F[3] = np.where(w > 0, w, 1.0)
    return F.astype(np.float32)


def test_spline_tdf_tube_axis_known_answer(orc):
    """On the axis of an ideal tube of wall radius 6 the 12 arms lie in the
    cross-section (e2, e3 perpendicular to the axis), every border field is
    perpendicular to the axis and opposite arms are equal: T = 1, and each
    border radius is the first sample past the wall (6.0 or 6.5 with step
    0.5).  The same in all three orientations."""
    for axis in (0, 1, 2):
        shape = [33, 33, 33]
        shape[axis] = 12
        t, r = orc.tdf_spline(_spline_tube(tuple(shape), 6.0, axis), 1.0, 10.0, 0.5)
        c = [16, 16, 16]
        c[axis] = 6
        assert t[tuple(c)] == 1.0, axis
        assert 6.0 <= r[tuple(c)] <= 6.5, axis


def test_spline_tdf_off_axis_and_invalid_cases(orc):
    """Off the axis the opposite arms differ, so T falls with the offset; a
    field pointing outward fails the border-direction test on the first arm
    (T = 0, radius = that arm's border, a sample just past the wall, / 12: the
    reference's partial mean); a
    uniform field has no border (T = 0, radius 0)."""
    F = _spline_tube((33, 33, 12), 6.0, 2)
    t, r = orc.tdf_spline(F, 1.0, 10.0, 0.5)
    ts = [float(t[16 + k, 16, 6]) for k in range(4)]
    assert all(a >= b for a, b in zip(ts, ts[1:])) and ts[3] < 1.0 and ts[3] > 0.2
    t, r = orc.tdf_spline(_spline_tube((33, 33, 12), 6.0, 2, sign=1.0), 1.0, 10.0, 0.5)
    assert t[16, 16, 6] == 0.0 and 6.0 / 12 <= r[16, 16, 6] <= 7.5 / 12
    u = np.zeros((4, 10, 10, 10), np.float32)
    u[0] = 0.3
    u[3] = 0.3
    t, r = orc.tdf_spline(u, 1.0, 5.0, 0.5)
    assert (t == 0).all() and (r == 0).all()


# ------------------------------------------------------------------ precision form: stochastic rounding
def test_half_conversion_matches_numpy(orc):
    """float -> half (round to nearest even) and back against numpy's IEEE
    conversions: random values over the whole half range (normal,
    subnormal, ties, overflow) and the special values."""
    g = rng(40)
    a = np.concatenate([
        g.uniform(-1, 1, 20000).astype(np.float32) * np.float32(2.0) ** g.integers(-30, 17, 20000).astype(np.float32),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, -65520, 1e-8, 2 ** -24, 2 ** -25, 3 * 2 ** -26, 2 ** -14,
                  np.inf, -np.inf, 1 + 2 ** -11, 1 + 3 * 2 ** -11], np.float32)])
    h = orc.half_from_float(a)
    with np.errstate(over="ignore"):
        ref = a.astype(np.float16).view(np.uint16)
    assert np.array_equal(h, ref)
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    back = orc.half_to_float(codes)
    fin = np.isfinite(codes.view(np.float16))
    assert np.array_equal(back[fin], codes.view(np.float16)[fin].astype(np.float32))


@pytest.mark.parametrize("storage", [1, 2])
def test_stochastic_rounding_is_one_of_the_neighbours_and_unbiased(orc, storage):
    g = rng(41)
    v = g.uniform(-3, 3, 4000).astype(np.float32) * np.float32(2 ** 14 if storage == 1 else 1)
    dec = orc.half_to_float if storage == 1 else (lambda h: (h.astype(np.uint32) << 16).view(np.float32))
    # r = 0 truncates towards zero; the magnitude rounds up when r makes the
    # discarded bits carry (r = 1023 for all but the last 2^-10 of an ulp)
    lo = dec(orc.s16_sr(v, np.zeros(v.size, np.uint32), storage)).astype(np.float64)
    up = lo
    if storage == 1:  # the next half away from zero
        up = np.nextafter(lo.astype(np.float16), (np.sign(v) * np.inf).astype(np.float16))
    else:  # the next bfloat16 away from zero
        b = (lo.astype(np.float32).view(np.uint32) + np.uint32(1 << 16)).view(np.float32)
        up = np.where(lo == v, lo, b)
    hi = np.where(lo == v, lo, up.astype(np.float64))
    assert np.all(np.abs(lo) <= np.abs(v)) and np.all(np.abs(hi) >= np.abs(v))
    # the mean over many random numbers is the value to within 2^-10 ulp
    # (the resolution of the 10 random bits) plus sampling error
    K = 2000
    r = g.integers(0, 1024, (K, v.size), dtype=np.uint32)
    got = dec(orc.s16_sr(np.broadcast_to(v, (K, v.size)).copy(), r, storage)).astype(np.float64)
    mean = got.mean(0)
    ulp = np.abs(hi - lo)
    assert np.all(np.abs(mean - v) <= 4 * ulp / np.sqrt(K) + ulp / 1024 + 1e-30)
    # each result is one of the two neighbours
    assert np.all((got == lo) | (got == hi))


def test_sr_hash_bits_are_uniform(orc):
    h = np.array([orc.sr_hash(i, 12345) for i in range(20000)], np.uint64)
    for b in range(32):
        frac = float(((h >> np.uint64(b)) & np.uint64(1)).mean())
        assert abs(frac - 0.5) < 0.02, (b, frac)
    assert orc.sr_hash(7, 1) != orc.sr_hash(7, 2) and orc.sr_hash(7, 1) == orc.sr_hash(7, 1)


def _f16_field(shape, seed):
    F = _field_from_phantom(shape, seed)
    return F, (F[:3] * np.float32(2 ** 14)).astype(np.float16).view(np.uint16)


def _field_from_phantom(shape, seed):
    import oracle as o
    from tsf import synth
    vol, _ = synth.tube_volume(shape, 4, 1.0, 3.0, seed, noise=0.05)
    return o.vector_field(o.blur(vol, 0.5, "f32"), 0.2, 1, "f32")


@pytest.mark.parametrize("shape", [(6, 48, 64), (4, 50, 66)])
def test_gvf_sr_one_launch_is_two_float_steps(orc, shape):
    """Two iterations are one launch with a float output and no 16-bit
    rounding of the iterate: equal to two FP64 Euler steps of the reference
    recurrence (mirror rule included) on the stored f, to float rounding."""
    F, f16 = _f16_field(shape, 42)
    out = orc.gvf_euler_sr16(f16, 1, 2.0 ** 14, 0.05, 2)
    f = np.concatenate([f16.view(np.float16).astype(np.float64) / 2 ** 14, np.ones((1,) + shape)])
    ref = orc.gvf_euler(f, 0.05, 2, "f64")
    err = np.abs(out[:3] - ref[:3]).max() / np.abs(ref[:3]).max()
    assert err < 2e-6, err
    assert np.allclose(out[3], np.sqrt((out[:3].astype(np.float64) ** 2).sum(0)), rtol=1e-6)


def test_gvf_low_memory_f32_equals_full_field(orc):
    """runLowMemoryGVF in float: each component alone with the stored |f|^2
    is the full-field recurrence exactly (the components couple only through
    |f|^2, computed in the same order)."""
    F = _field_from_phantom((9, 14, 17), 48)
    for iters in (0, 1, 6):
        lo = orc.gvf_euler_low_memory(F, 0.05, iters)
        full = orc.gvf_euler(F, 0.05, iters, "f32")
        assert np.array_equal(lo, full), iters


def test_gvf_low_memory_sn16_rounds_only_the_magnitude(orc):
    """In the 16-bit mode |f|^2 is stored as SNORM16 once: on a field whose
    |f|^2 is exactly 0 or 1 (one component +-1 per voxel, or none) that
    rounding is exact and the one-component GVF equals the full-field 16-bit
    GVF bit for bit; on a general field it differs, by less than a few
    codes."""
    g = rng(49)
    shape = (8, 11, 13)
    F = np.zeros((4,) + shape, np.float32)
    comp = g.integers(0, 4, shape)
    sign = np.where(g.random(shape) < 0.5, -1.0, 1.0).astype(np.float32)
    for c in range(3):
        F[c][comp == c] = sign[comp == c]
    F[3] = 1.0
    for iters in (1, 5):
        assert np.array_equal(orc.gvf_euler_low_memory(F, 0.05, iters, sn16=True), orc.gvf_euler_sn16(F, 0.05, iters))
    Fs = orc.sn16_round(_field_from_phantom(shape, 50))
    lo = orc.gvf_euler_low_memory(Fs, 0.05, 6, sn16=True)
    full = orc.gvf_euler_sn16(Fs, 0.05, 6)
    assert not np.array_equal(lo[:3], full[:3])
    assert np.abs(lo[:3] - full[:3]).max() <= 8 / 32767


def _sn16_field(shape, seed):
    import oracle as o
    F = _field_from_phantom(shape, seed)
    return F, o.sn16_store(F[:3]).view(np.uint16)


def _sn16_values(codes):
    return np.maximum(codes.view(np.int16).astype(np.float64), -32767.0) / 32767.0


@pytest.mark.parametrize("shape", [(6, 48, 64), (4, 50, 66)])
def test_gvf_sr_snorm16_f_one_launch_is_two_float_steps(orc, shape):
    """f in SNORM16 (the iterate in F16 at unit 2^14): one launch from v0 = f
    is two FP64 Euler steps on f's SNORM values max(s, -32767) / 32767, to
    float rounding (the decode as s * (2^14 / 32767) in float adds an ulp)."""
    F, codes = _sn16_field(shape, 45)
    out = orc.gvf_euler_sr16(codes, 3, 0.0, 0.05, 2)
    f = np.concatenate([_sn16_values(codes), np.ones((1,) + shape)])
    ref = orc.gvf_euler(f, 0.05, 2, "f64")
    err = np.abs(out[:3] - ref[:3]).max() / np.abs(ref[:3]).max()
    assert err < 2e-6, err


def test_gvf_sr_snorm16_f_odd_step_rounds_to_nearest_half(orc):
    """One step (an odd count) from SNORM16 f: the FP64 step rounded to the
    nearest F16 at unit 2^14, to within half a half-ulp plus float error."""
    shape = (5, 48, 64)
    F, codes = _sn16_field(shape, 46)
    out = orc.gvf_euler_sr16(codes, 3, 0.0, 0.05, 1)[:3].astype(np.float64)
    f = np.concatenate([_sn16_values(codes), np.ones((1,) + shape)])
    ref = orc.gvf_euler(f, 0.05, 1, "f64")[:3]
    half_ulp = 2.0 ** -11 * np.abs(ref) + 2.0 ** -39
    assert np.all(np.abs(out - ref) <= half_ulp * 1.001 + 1e-7 * np.abs(ref).max())
    # and it is a half at unit 2^14
    assert np.array_equal((out * 2 ** 14).astype(np.float16).astype(np.float64) / 2 ** 14, out)


def test_gvf_sr_snorm16_f_is_closer_to_the_unrounded_field(orc):
    """Against FP64 GVF of the float vector field itself (not of a stored
    copy), 200 steps: the form with f in SNORM16 lands several times closer
    than the one with f in F16, since f's rounding dominates the error and
    SNORM16 resolves components in [-1, 1] to 1.5e-5 where F16 has 2^-12
    relative."""
    shape = (16, 48, 64)
    F, codes = _sn16_field(shape, 47)
    f16 = (F[:3] * np.float32(2 ** 14)).astype(np.float16).view(np.uint16)
    ref = orc.gvf_euler(F.astype(np.float64), 0.05, 200, "f64")[:3]
    e_sn = orc.gvf_euler_sr16(codes, 3, 0.0, 0.05, 200, seed=5)[:3]
    e_h = orc.gvf_euler_sr16(f16, 1, 2.0 ** 14, 0.05, 200, seed=5)[:3]
    r = lambda a: np.linalg.norm(a - ref) / np.linalg.norm(ref)
    print(f"200 steps vs FP64 of the float field: f SNORM16 {r(e_sn):.2e}, f F16 {r(e_h):.2e}")
    assert r(e_sn) < 0.6 * r(e_h)


def test_gvf_sr_mu_zero_keeps_f(orc):
    """mu = 0 with v0 = f: every update of an interior voxel is
    v (1 - |f|^2) + f |f|^2 = f up to float rounding, so after six steps
    (two stochastic roundings) each voxel is within one 16-bit ulp of f;
    face voxels take their mirror's (0 -> 2, N-1 -> N-3)."""
    shape = (5, 48, 64)
    F, f16 = _f16_field(shape, 43)
    out = orc.gvf_euler_sr16(f16, 1, 2.0 ** 14, 0.0, 6)
    f = f16.view(np.float16).astype(np.float32) / np.float32(2 ** 14)
    dev = np.abs(out[:3, 1:-1, 1:-1, 1:-1] - f[:, 1:-1, 1:-1, 1:-1])
    assert np.all(dev <= np.abs(f[:, 1:-1, 1:-1, 1:-1]) * 2.0 ** -10 + 2.0 ** -38)
    rm = lambda n: np.r_[2, np.arange(1, n - 1), n - 3]
    Z, Y, X = shape
    assert np.array_equal(out[:3], out[:3][:, rm(Z)][:, :, rm(Y)][:, :, :, rm(X)])


def test_gvf_sr_converges_where_round_to_nearest_stalls(orc):
    """Against the FP64 iterate after 200 steps, stochastic rounding of the
    stored iterate stays within a few 16-bit ulps in L2 while round to
    nearest (tsf_gvf_euler's F16 storage, every launch rounded to nearest)
    stalls further away; both start from the same stored f."""
    shape = (16, 48, 64)
    F, f16 = _f16_field(shape, 44)
    f = np.concatenate([f16.view(np.float16).astype(np.float64) / 2 ** 14, np.ones((1,) + shape)])
    ref = orc.gvf_euler(f, 0.05, 200, "f64")[:3]
    sr = orc.gvf_euler_sr16(f16, 1, 2.0 ** 14, 0.05, 200, seed=5)[:3]
    f32 = f[:3].astype(np.float32)
    rne = f32.copy()
    for _ in range(100):  # two float steps, then round to nearest
        rne = orc.f16_round(orc.gvf_step(f32, orc.gvf_step(f32, rne, 0.05), 0.05))
    e_sr = np.linalg.norm(sr - ref) / np.linalg.norm(ref)
    e_rne = np.linalg.norm(rne - ref) / np.linalg.norm(ref)
    print(f"200 steps: stochastic {e_sr:.2e}, nearest {e_rne:.2e}")
    assert e_sr < 3e-4 and e_sr < 0.6 * e_rne
