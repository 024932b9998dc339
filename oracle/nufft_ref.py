"""CPU (numpy, FP64) restatement of CuRIG's improved-w-stacking gridder.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` may import this module, and only as the
checker.  It is sized for small problems (the full 3-D grid is held in memory).

Every function restates one step of the reference's algorithm and cites the
reference file:line it follows (paths relative to the CuRIG tree).  Where the
reference has undefined behaviour, this restatement follows the direct-sum DFT
definition and says so:

* coordinates beyond +-3pi index out of bounds in the reference
  (src/fourier/conv.cu:213-215, single periodic wrap); here the grid index is
  taken modulo nf, which equals the single wrap wherever the reference is
  defined and equals the DFT's 2pi periodicity everywhere;
* when ceil(t - kw/2) == t - kw/2 the reference loop visits kw+1 cells and reads
  one uninitialised kernel value (src/fourier/conv.cu:204-210 vs :21); here
  exactly kw cells are visited (the extra cell's kernel value is 0).

Pinning (tests/test_oracle.py, tests/test_plan_math.py): the whole pipeline
against the FP64 direct-sum DFT (oracle/dft.c) at the requested epsilon;
exact adjointness (reference assertion python/curagridder/test_wst.py:77);
kw/beta table, next235beven by brute force, Gauss-Legendre nodes against
numpy.polynomial.legendre.leggauss, the kernel's Fourier series against
scipy.integrate.quad, the FFT step against a direct sum on tiny planes.
"""
import math

import numpy as np

SPEEDOFLIGHT = 299792458.0          # include/utils.h:70
EPSILON = 1.1e-16                   # include/utils.h:66
MAX_KERNEL_WIDTH = 28               # include/curafft_plan.h:14
M_1_2PI = 0.159154943091895336      # include/utils.h:61


# ----------------------------------------------------------------------------
# host plan math
# ----------------------------------------------------------------------------
def setup_conv_opts(eps, upsampfac):
    """kw, beta, c = 4/kw^2 and the error code.

    src/fourier/conv_interp_invoker.cu:46-116.  Returns (ier, kw, beta, c).
    ier: 0 ok, 1 eps clamped to 1.1e-16, 2 invalid sigma (sigma <= 1).
    """
    if upsampfac != 2.0 and upsampfac <= 1.0:
        return 2, 0, 0.0, 0.0
    ier = 0
    if eps < EPSILON:
        eps = EPSILON
        ier = 1
    kw = int(math.ceil(-math.log10(eps / 10.0)))
    if upsampfac != 2.0:
        kw = int(math.ceil(-math.log(eps) / (math.pi * math.sqrt(1.0 - 1.0 / upsampfac))))
    kw = max(2, kw)
    kw = min(kw, MAX_KERNEL_WIDTH)
    betaoverns = 2.30
    if kw == 2:
        betaoverns = 2.20
    if kw == 3:
        betaoverns = 2.26
    if kw == 4:
        betaoverns = 2.38
    if upsampfac != 2.0:
        betaoverns = 0.97 * math.pi * (1 - 1 / (2 * upsampfac))
    beta = betaoverns * kw
    c = 4.0 / (kw * kw)
    return ier, kw, beta, c


def next235beven(n, b=1):
    """Smallest even n' >= n with prime factors <= 5 and n' % b == 0.

    src/utils.cu:251-276 (next235even, contrib/common.cpp:15-32, is b=1)."""
    if n <= 2:
        return 2
    if n % 2 == 1:
        n += 1
    nplus = n - 2
    numdiv = 2
    while numdiv > 1 or nplus % b != 0:
        nplus += 2
        numdiv = nplus
        while numdiv % 2 == 0:
            numdiv //= 2
        while numdiv % 3 == 0:
            numdiv //= 3
        while numdiv % 5 == 0:
            numdiv //= 5
    return nplus


def get_num_cells(ms, kw, upsampfac):
    """src/fourier/conv_interp_invoker.cu:29-44."""
    nf = int(upsampfac * ms)
    if nf < 2 * kw:
        nf = 2 * kw
    return next235beven(nf, 1)


def set_nhg_type3(S, X, kw, upsampfac):
    """Number of w planes nf, grid spacing h, scale gamma.  contrib/common.cpp:161-196."""
    nss = kw + 1
    Xsafe, Ssafe = X, S
    if X == 0.0:
        if S == 0.0:
            Xsafe, Ssafe = 1.0, 1.0
        else:
            Xsafe = max(Xsafe, 1.0 / S)
    else:
        Ssafe = max(Ssafe, 1.0 / X)
    nfd = 2.0 * upsampfac * Ssafe * Xsafe / math.pi + nss
    if not math.isfinite(nfd):
        nfd = 0.0
    nf = int(nfd)
    if nf < 2 * kw:
        nf = 2 * kw
    nf = next235beven(nf, 1)
    h = 2 * math.pi / nf
    gam = nf / (2.0 * upsampfac * Ssafe)
    return nf, h, gam


def gauss_legendre(n):
    """n-point Gauss-Legendre nodes (ascending) and weights on [-1, 1].

    The reference computes the same rule with the Glaser-Liu-Rokhlin method
    (contrib/legendre_rule_fast.c:15-92); the rule is unique, so it is restated
    here by Newton iteration on the three-term recurrence."""
    x = np.empty(n)
    w = np.empty(n)
    for i in range(n):
        z = math.cos(math.pi * (i + 0.75) / (n + 0.5))
        for _ in range(100):
            p0, p1 = 1.0, z
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
            dp = n * (z * p1 - p0) / (z * z - 1.0)
            dz = p1 / dp
            z -= dz
            if abs(dz) < 1e-16:
                break
        p0, p1 = 1.0, z
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        dp = n * (z * p1 - p0) / (z * z - 1.0)
        x[n - 1 - i] = z
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp)
    return x, w


def es_kernel(d, kw, beta, c):
    """phi(d) = exp(beta sqrt(1 - c d^2)) for |d| < kw/2 else 0.  src/fourier/conv.cu:17-25."""
    d = np.asarray(d, dtype=np.float64)
    out = np.zeros_like(d)
    m = np.abs(d) < kw / 2.0
    out[m] = np.exp(beta * np.sqrt(1.0 - c * d[m] * d[m]))
    return out


def _quadrature(kw, beta, c):
    """Scaled nodes alpha*t_n and g_n = alpha*w_n*phi(alpha*t_n) over all 2p nodes.

    src/fourier/deconv.cu:72-86 (p = (int)(2 + 3*alpha), 2p-point rule; the
    reference sums the p nodes of one half and doubles, which for the even
    integrand equals the full 2p-point sum used here)."""
    alpha = kw / 2.0
    p = int(2 + 3.0 * alpha)
    t, wq = gauss_legendre(2 * p)
    x = t * alpha
    phi = np.where(np.abs(x) <= kw / 2.0, np.exp(beta * np.sqrt(np.maximum(1.0 - c * x * x, 0.0))), 0.0)
    g = alpha * wq * phi
    return x, g


def fseries_grid(nf, kw, beta, c):
    """phihat(k) * (-1)^k for k = 0..nf/2: the 1-D correction factors.

    src/fourier/deconv.cu:13-24 with N = nf/2+1: cos((x_n - nf/2)/(nf/2) pi k)
    = (-1)^k cos(2 pi k x_n / nf)."""
    x, g = _quadrature(kw, beta, c)
    k = np.arange(nf // 2 + 1, dtype=np.float64)
    half = nf // 2
    return np.array([np.sum(g * np.cos((x - half) / half * math.pi * kk)) for kk in k])


def fseries_at(z, kw, beta, c):
    """phihat(z) = int phi(x) cos(x z) dx at arbitrary z.  src/fourier/deconv.cu:45-58."""
    x, g = _quadrature(kw, beta, c)
    z = np.asarray(z, dtype=np.float64)
    return np.sum(g[None, :] * np.cos(np.outer(z.ravel(), x)), axis=1).reshape(z.shape)


def plan_params(nrow, nx, ny, fov, freq, uvw, eps, sigma, sign):
    """Everything gridder_setting computes.  src/astro/cugridder.cu:29-101, 105-260.

    Image convention: dirty[nx][ny]; N1 = ny (fast, v axis, nf1), N2 = nx (u axis, nf2);
    the plan's x axis is v and its y axis is u (setup_plan(..., d_v, d_u, ...), :188)."""
    ier, kw, beta, c = setup_conv_opts(eps, sigma)
    N1, N2 = ny, nx
    f_over_c = freq / SPEEDOFLIGHT
    u = uvw[:, 0] * f_over_c * 2 * math.pi
    v = uvw[:, 1] * f_over_c * 2 * math.pi
    w = uvw[:, 2] * (f_over_c * sign) * 2 * math.pi
    pixelsize_x = fov / 180.0 * math.pi / N2
    pixelsize_y = fov / 180.0 * math.pi / N1
    l_min = 0 - 0.5 * pixelsize_x * N2
    m_min = 0 - 0.5 * pixelsize_y * N1
    n_lm = math.sqrt(1.0 - l_min ** 2 - m_min ** 2)
    if nrow > 0:
        i_max, i_min = float(np.max(w)), float(np.min(w))
    else:
        i_max = i_min = 0.0
    i_center = (i_max + i_min) / 2.0
    i_half_width = (i_max - i_min) / 2.0
    o_min = n_lm - 1
    o_center = o_min / 2.0
    o_half_width = abs(o_min / 2.0)
    nf3, h, gamma = set_nhg_type3(o_half_width, i_half_width, kw, sigma)
    nf1 = get_num_cells(N1, kw, sigma)
    nf2 = get_num_cells(N2, kw, sigma)
    return dict(ier=ier, kw=kw, beta=beta, c=c, N1=N1, N2=N2, nf1=nf1, nf2=nf2, nf3=nf3,
                h=h, gamma=gamma, i_center=i_center, i_half_width=i_half_width,
                o_center=o_center, o_half_width=o_half_width,
                pixelsize_x=pixelsize_x, pixelsize_y=pixelsize_y,
                xu=u * pixelsize_x, yv=v * pixelsize_y, w=w)


def _k_quadrant(N1, N2, pixelsize_x, pixelsize_y):
    """z = n - 1 on the (N2/2+1) x (N1/2+1) quadrant.  src/astro/precomp.cu:212-222."""
    row = np.arange(N2 // 2 + 1, dtype=np.float64)[:, None]
    col = np.arange(N1 // 2 + 1, dtype=np.float64)[None, :]
    return np.sqrt(1.0 - (row * pixelsize_x) ** 2 - (col * pixelsize_y) ** 2) - 1


def _grid_coords(t, kw, nf):
    """start index and kw kernel offsets for rescaled coordinate t (grid units)."""
    start = np.ceil(t - kw / 2.0).astype(np.int64)
    offs = start[:, None] + np.arange(kw)[None, :] - t[:, None]
    idx = np.mod(start[:, None] + np.arange(kw)[None, :], nf)
    return idx, offs


def _rescale(x, nf):
    """SHIFT_RESCALE(x, N, pirange=1) = (x/2pi + 0.5) N.  include/conv.h:21.
    Coordinates are first folded into [-pi, pi) (DFT periodicity)."""
    xf = x - np.floor(x / (2 * math.pi) + 0.5) * (2 * math.pi)
    return (xf * M_1_2PI + 0.5) * nf


def _pixel_maps(N1, N2, nf1, nf2):
    """FFT-ordered cell of each output pixel and the quadrant index (CMCL mode).
    src/fourier/cuft.cu:536-552, src/fourier/deconv.cu:246-263."""
    idx = np.arange(N1 * N2)
    row = idx // N1
    col = idx % N1
    w1 = np.where(col >= N1 // 2, col - N1 // 2, nf1 + col - N1 // 2)
    w2 = np.where(row >= N2 // 2, row - N2 // 2, nf2 + row - N2 // 2)
    idx_z = np.abs(col - N1 // 2) + np.abs(row - N2 // 2) * (N1 // 2 + 1)
    return row, col, w1, w2, idx_z


def ms2dirty(uvw, vis, wgt, nx, ny, fov, freq, eps, sigma=2.0, sign=-1):
    """Reference ms2dirty pipeline, FP64.  Returns complex128 dirty[nx][ny].

    src/astro/ra_exec.cu:141-308 (exec_vis2dirty) via cugridder.cu:330-453."""
    uvw = np.asarray(uvw, dtype=np.float64)
    nrow = uvw.shape[0]
    pp = plan_params(nrow, nx, ny, fov, freq, uvw, eps, sigma, sign)
    kw, beta, c = pp["kw"], pp["beta"], pp["c"]
    N1, N2, nf1, nf2, nf3 = pp["N1"], pp["N2"], pp["nf1"], pp["nf2"], pp["nf3"]
    C, D, gamma, h = pp["i_center"], pp["o_center"], pp["gamma"], pp["h"]
    cvis = np.asarray(vis, dtype=np.complex128).ravel().copy()
    # cura_prestage: vis *= wgt (ra_exec.cu:98-105); c *= e^{i D w} (cuft.cu:19-39)
    if wgt is not None:
        cvis = cvis * np.asarray(wgt, dtype=np.float64).ravel()
    cvis = cvis * np.exp(1j * D * pp["w"])
    wp = (pp["w"] - C) / gamma                              # pre_stage_2, cuft.cu:69
    z = (_k_quadrant(N1, N2, pp["pixelsize_x"], pp["pixelsize_y"]) - D) * gamma * h  # cuft.cu:73
    fk1 = fseries_grid(nf1, kw, beta, c)
    fk2 = fseries_grid(nf2, kw, beta, c)
    fk3 = fseries_at(z, kw, beta, c).ravel()
    # conv_3d_nputsdriven: plan x = v (nf1), y = u (nf2), z = w' (nf3)   conv.cu:144-228
    ix, kx = _grid_coords(_rescale(pp["yv"], nf1), kw, nf1)
    iy, ky = _grid_coords(_rescale(pp["xu"], nf2), kw, nf2)
    iz, kz = _grid_coords((wp * M_1_2PI + 0.5) * nf3, kw, nf3)
    phx, phy, phz = es_kernel(kx, kw, beta, c), es_kernel(ky, kw, beta, c), es_kernel(kz, kw, beta, c)
    fw = np.zeros((nf3, nf2, nf1), dtype=np.complex128)
    if nrow > 0:
        lin = (iz[:, :, None, None] * nf2 + iy[:, None, :, None]) * nf1 + ix[:, None, None, :]
        val = cvis[:, None, None, None] * phz[:, :, None, None] * phy[:, None, :, None] * phx[:, None, None, :]
        np.add.at(fw.reshape(-1), lin.reshape(-1), val.reshape(-1))
    # cura_cufft: batched 2-D FFT, sign +1, unnormalised   cuft.cu:127-143
    fw = np.fft.ifft2(fw, axes=(1, 2)) * (nf1 * nf2)
    # w_term_dft: sum_p fw[p] e^{+i z'(p - nf3/2)}   cuft.cu:520-572
    row, col, w1, w2, idx_z = _pixel_maps(N1, N2, nf1, nf2)
    planes = fw[:, w2, w1]                                   # [nf3][N]
    zz = z.ravel()[idx_z]
    ph = np.exp(1j * np.outer(np.arange(nf3) - nf3 // 2, zz))
    fk = np.sum(planes * ph, axis=0)
    # deconv_2d type 1 (mode_flag 1)   deconv.cu:236-276
    fk = fk / (fk1[np.abs(col - N1 // 2)] * fk2[np.abs(row - N2 // 2)])
    # w_term_deconv   deconv.cu:371-399 (flag = iflag = +1)
    px, py = pp["pixelsize_x"], pp["pixelsize_y"]
    nlm = np.sqrt(1.0 - ((row - N2 // 2) * px) ** 2 - ((col - N1 // 2) * py) ** 2)
    phase = ((nlm - 1) - D) * C * 1
    fk = fk * np.exp(1j * phase) / fk3[idx_z]
    # div_n_lm   ra_exec.cu:22-37
    fk = fk / nlm
    return fk.reshape(nx, ny)


def dirty2ms(uvw, dirty, wgt, nx, ny, fov, freq, eps, sigma=2.0, sign=-1):
    """Reference dirty2ms pipeline, FP64.  Returns complex128 vis[nrow].

    src/astro/ra_exec.cu:325-436 (exec_dirty2vis) via cugridder.cu:455-582."""
    uvw = np.asarray(uvw, dtype=np.float64)
    nrow = uvw.shape[0]
    pp = plan_params(nrow, nx, ny, fov, freq, uvw, eps, sigma, sign)
    kw, beta, c = pp["kw"], pp["beta"], pp["c"]
    N1, N2, nf1, nf2, nf3 = pp["N1"], pp["N2"], pp["nf1"], pp["nf2"], pp["nf3"]
    C, D, gamma, h = pp["i_center"], pp["o_center"], pp["gamma"], pp["h"]
    wp = (pp["w"] - C) / gamma
    z = (_k_quadrant(N1, N2, pp["pixelsize_x"], pp["pixelsize_y"]) - D) * gamma * h
    fk1 = fseries_grid(nf1, kw, beta, c)
    fk2 = fseries_grid(nf2, kw, beta, c)
    fk3 = fseries_at(z, kw, beta, c).ravel()
    row, col, w1, w2, idx_z = _pixel_maps(N1, N2, nf1, nf2)
    px, py = pp["pixelsize_x"], pp["pixelsize_y"]
    nlm = np.sqrt(1.0 - ((row - N2 // 2) * px) ** 2 - ((col - N1 // 2) * py) ** 2)
    fk = np.asarray(dirty, dtype=np.complex128).reshape(-1).copy()
    fk = fk / nlm                                            # cura_rscaling  ra_exec.cu:332
    fk = fk * np.exp(1j * (((nlm - 1) - D) * C * -1)) / fk3[idx_z]   # w_term_deconv flag -1
    fw0 = fk / (fk1[np.abs(col - N1 // 2)] * fk2[np.abs(row - N2 // 2)])  # deconv_2d type 2
    # w_term_idft: plane p = fw0 e^{-i z'(p - nf3/2)}   cuft.cu:575-612
    fw = np.zeros((nf3, nf2, nf1), dtype=np.complex128)
    zz = z.ravel()[idx_z]
    fw[:, w2, w1] = fw0[None, :] * np.exp(-1j * np.outer(np.arange(nf3) - nf3 // 2, zz))
    # cura_cufft sign -1 (forward)
    fw = np.fft.fft2(fw, axes=(1, 2))
    # interp_3d_nputsdriven   interp.cu:143-228
    ix, kx = _grid_coords(_rescale(pp["yv"], nf1), kw, nf1)
    iy, ky = _grid_coords(_rescale(pp["xu"], nf2), kw, nf2)
    iz, kz = _grid_coords((wp * M_1_2PI + 0.5) * nf3, kw, nf3)
    phx, phy, phz = es_kernel(kx, kw, beta, c), es_kernel(ky, kw, beta, c), es_kernel(kz, kw, beta, c)
    out = np.zeros(nrow, dtype=np.complex128)
    if nrow > 0:
        flat = fw.reshape(-1)
        lin = (iz[:, :, None, None] * nf2 + iy[:, None, :, None]) * nf1 + ix[:, None, None, :]
        ker = phz[:, :, None, None] * phy[:, None, :, None] * phx[:, None, None, :]
        out = np.sum((flat[lin] * ker).reshape(nrow, -1), axis=1)
    # cura_fw: vis *= wgt; shift_corr e^{-i D w}   ra_exec.cu:73-89, :56-65
    if wgt is not None:
        out = out * np.asarray(wgt, dtype=np.float64).ravel()
    out = out * np.exp(1j * ((wp * gamma + C) * D * -1))
    return out


# ----------------------------------------------------------------------------
# generic NUFFT types 1 and 2 (direct sums; small sizes)
# ----------------------------------------------------------------------------
def nufft_direct(type_, N, coords, data, iflag):
    """Direct sums of the generic NUFFT, as the reference's unit drivers define
    their ground truth (test/nufft_1d_test.cu:289-296, test/nufft_2d_1_test.cu
    crude direct sum, test/nufft_2d_2_test.cu:319-328):

      type 1: fk[k3][k2][k1] = sum_j c_j e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}
      type 2: c_j = sum_k fk[k3][k2][k1] e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}

    with k' = k - N/2 (integer halving), N = (N1, [N2, [N3]]) and coords =
    (u, [v, [w]]) in radians.  fk has shape N[::-1] (N1 the fast index)."""
    N = tuple(int(n) for n in N)
    dim = len(N)
    assert dim == len(coords) and 1 <= dim <= 3
    E = []
    for n, x in zip(N, coords):
        k = np.arange(n) - n // 2
        E.append(np.exp(1j * iflag * np.outer(np.asarray(x, dtype=np.float64), k)))
    spec = {1: ("ja->a", "a,ja->j"), 2: ("ja,jb->ba", "ba,ja,jb->j"),
            3: ("ja,jb,jc->cba", "cba,ja,jb,jc->j")}[dim]
    if type_ == 1:
        c = np.asarray(data, dtype=np.complex128)
        return np.einsum("j," + spec[0], c, *E)
    fk = np.asarray(data, dtype=np.complex128).reshape(N[::-1])
    return np.einsum(spec[1], fk, *E)


def nufft3_direct(x, s, data, iflag, direction):
    """Direct sums of the 1-D type-3 transform with inputs x (M) and outputs s
    (N), as the reference's drivers define their ground truth
    (test/nufft_1d_3_1_test.cu:210-219, direction 1 with iflag +1;
    test/nufft_1d_3_2_test.cu:180-187, direction 0 with iflag -1):

      direction 1: f_i = sum_j c_j e^{i iflag s_i x_j}   (data = c, returns f)
      direction 0: c_j = sum_i f_i e^{i iflag s_i x_j}   (data = f, returns c)"""
    E = np.exp(1j * iflag * np.outer(np.asarray(s, dtype=np.float64), np.asarray(x, dtype=np.float64)))  # [N, M]
    d = np.asarray(data, dtype=np.complex128)
    return E @ d if direction == 1 else d @ E
