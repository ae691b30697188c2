"""CPU oracle of the path-guiding model -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this module; the library never
does. Everything is float64 numpy, with optional emulation of the GPU's operand rounding (FP16 grid
features, BF16 MLP operands and activations, FP32 position arithmetic) so that the GPU result can be
bounded instead of merely compared with a loose tolerance.

What it restates:
  * vMF lobes and mixtures: VMF::eval / VMF::sample / max_kappa (vmf.h:10-114), WVMF + sample_vmm
    (vmf.h:117-127) with sample_small_distrib (rng.h:437-470) and sample_uniform_sphere (rng.h:284-292),
    orthonormal_basis (maths.h:814-821).
  * Adam: adam_optimizer_step (gpu/shaders/adam.slang:25-51), lerp form with bias correction.
  * The network itself (hash grid, MLP, mixture head, loss) has no reference counterpart: the reference
    fits vMF mixtures with EM (vmf.cpp) and has no learned guide. Its definition here is the
    specification the GPU kernels implement:
      - position -> u = (x - lo) * inv_ext (FP32), per level l: resolution res_l, cell = min(floor(u *
        res_l), res_l - 1), trilinear weights in FP32, corner index dense x + s(y + s z) with s = res_l + 1
        when s^3 <= T, else the spatial hash of Mueller et al. 2022 (x ^ y*2654435761 ^ z*805459861) mod T;
      - input row [32 features, normal, wo, 1, 0 x 25] -> ReLU(W1) -> ReLU(W2) -> W3 -> 48 raw outputs;
      - head: lobe k uses outputs 5k..5k+4 = (logit, mean xyz, log kappa); softmax weights, normalised
        means, kappa = clamp(exp(.), 1e-2, 1e4);
      - loss sum_i -c_i log q(wi_i);
      - FP8 inference (forward_fp8): each layer input x -> e4m3(x / s_in) with s_in a per-tensor power
        of two from the recorded activation maximum (one power of two of headroom), W -> e4m3(W / s_w)
        with s_w the smallest power of two bringing max |W| to 448, FP32 sums scaled back by s_in s_w.
"""

import math

import numpy as np

LEVELS = 16
FEATURES = 2
LOBES = 8
N_IN = 64
N_HID = 64
N_OUT = 48
MAX_KAPPA = 1e4  # VMF::max_kappa (vmf.h:110)
MIN_KAPPA = 1e-2
W1_SIZE = 64 * 64
W2_SIZE = 64 * 64
W3_SIZE = 64 * N_OUT
MLP_SIZE = W1_SIZE + W2_SIZE + W3_SIZE
F32 = np.float32
U32 = np.uint32


# ---------------------------------------------------------------------------------------------------
# rounding helpers
# ---------------------------------------------------------------------------------------------------
def bf16(a):
    """Round to BF16 (nearest even) through FP32, as the GPU does; returns float64 values."""
    f = np.asarray(a, dtype=np.float64).astype(F32)
    u = f.view(U32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16) & 0xFFFF
    return (r.astype(U32) << U32(16)).view(F32).astype(np.float64)


def fp16(a):
    """Round FP32 values to FP16 (nearest even); returns float64 values."""
    return np.asarray(a, dtype=F32).astype(np.float16).astype(np.float64)


# ---------------------------------------------------------------------------------------------------
# grid layout
# ---------------------------------------------------------------------------------------------------
def resolutions(base=16.0, max_res=256.0):
    """floor(base * b^l) with b = (max / base)^(1 / (L - 1)), in float64."""
    g = math.exp(math.log(float(F32(max_res)) / float(F32(base))) / (LEVELS - 1))
    return [int(math.floor(float(F32(base)) * g**l + 1e-9)) for l in range(LEVELS)]


def normalisation(bound_lo, bound_hi):
    """Scene bound padded by 1% (at least 1e-4) on each side, FP32 arithmetic: (lo, inv_ext)."""
    lo = np.asarray(bound_lo, dtype=F32)
    hi = np.asarray(bound_hi, dtype=F32)
    ext = hi - lo
    pad = np.maximum(ext * F32(0.01), F32(1e-4))
    return lo - pad, F32(1.0) / (ext + F32(2.0) * pad)


def grid_index(x, y, z, res, log2_t):
    t = 1 << log2_t
    side = res + 1
    x, y, z = (np.asarray(v, dtype=U32) for v in (x, y, z))
    if side**3 <= t:
        return x + U32(side) * (y + U32(side) * z)
    return (x ^ (y * U32(2654435761)) ^ (z * U32(805459861))) & U32(t - 1)


class Layout:
    """Grid resolutions, table size and the position normalisation of one guide."""

    def __init__(self, log2_t=19, base=16.0, max_res=256.0, bound_lo=(0, 0, 0), bound_hi=(1, 1, 1)):
        self.log2_t = log2_t
        self.res = resolutions(base, max_res)
        self.lo, self.inv_ext = normalisation(bound_lo, bound_hi)
        self.n_grid = LEVELS * (1 << log2_t) * FEATURES

    def corners(self, x):
        """Per level: (idx (n, 8) into the table of that level, weights (n, 8) float64 of FP32 values)."""
        u = (np.asarray(x, dtype=F32) - self.lo) * self.inv_ext
        u = np.clip(u, F32(0), F32(1)).astype(F32)
        out = []
        for l, r in enumerate(self.res):
            rf = F32(r)
            xs = (u * rf).astype(F32)
            fl = np.minimum(np.floor(xs), rf - F32(1)).astype(F32)
            cell = fl.astype(U32)
            p = (xs - fl).astype(F32)
            q = (F32(1) - p).astype(F32)
            idx = np.empty((len(u), 8), dtype=np.int64)
            w = np.empty((len(u), 8), dtype=np.float64)
            for k in range(8):
                d = (k & 1, (k >> 1) & 1, (k >> 2) & 1)
                idx[:, k] = grid_index(cell[:, 0] + U32(d[0]), cell[:, 1] + U32(d[1]), cell[:, 2] + U32(d[2]), r,
                                       self.log2_t).astype(np.int64) + (l << self.log2_t)
                wk = (p[:, 0] if d[0] else q[:, 0]) * (p[:, 1] if d[1] else q[:, 1])
                wk = (wk.astype(F32) * (p[:, 2] if d[2] else q[:, 2])).astype(F32)
                w[:, k] = wk
            out.append((idx, w))
        return out


# ---------------------------------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------------------------------
def split_params(params, layout):
    p = np.asarray(params, dtype=np.float64)
    g = p[: layout.n_grid].reshape(-1, FEATURES)
    o = layout.n_grid
    w1 = p[o:o + W1_SIZE].reshape(64, 64)
    w2 = p[o + W1_SIZE:o + W1_SIZE + W2_SIZE].reshape(64, 64)
    w3 = p[o + W1_SIZE + W2_SIZE:o + MLP_SIZE].reshape(64, N_OUT)
    return g, w1, w2, w3


class Net:
    """Parameters of one guide in the precision a forward pass uses (quantised=True: FP16 grid, BF16
    weights, as the GPU holds them; quantised=False: the FP32 master values)."""

    def __init__(self, params, layout, quantised=True):
        g, w1, w2, w3 = split_params(params, layout)
        self.layout = layout
        self.quantised = quantised
        if quantised:
            self.grid, self.w = fp16(g), [bf16(w1), bf16(w2), bf16(w3)]
        else:
            self.grid, self.w = g.copy(), [w1.copy(), w2.copy(), w3.copy()]


def encode(net, queries):
    """Input rows (n, 64) in float64 (grid features summed exactly) plus the per-feature FP32 summation
    error bound of the GPU (n, 64)."""
    q = np.asarray(queries, dtype=F32).reshape(-1, 9)
    n = len(q)
    x = np.zeros((n, N_IN))
    err = np.zeros((n, N_IN))
    for l, (idx, w) in enumerate(net.layout.corners(q[:, 0:3])):
        for f in range(FEATURES):
            terms = w * net.grid[idx, f]
            x[:, 2 * l + f] = terms.sum(axis=1)
            err[:, 2 * l + f] = 2.0 * 8 * 2.0**-24 * np.abs(terms).sum(axis=1)
    x[:, 32:38] = q[:, 3:9]
    x[:, 38] = 1.0
    return x, err


def _round_interval(v, d):
    """BF16 of v and the radius of {bf16(t) : |t - v| <= d} around it (rounding is monotone)."""
    c = bf16(v)
    return c, np.maximum(bf16(v + d) - c, c - bf16(v - d))


def forward(net, queries, with_bound=False):
    """Raw outputs (n, 48). With quantised parameters the activations are rounded to BF16 between layers
    as on the GPU; with_bound also returns a bound on |GPU - oracle| per output: an interval carried
    through the FP32 accumulation error (2 * 64 * 2^-24 of the absolute sum) and every BF16 rounding."""
    x, ex = encode(net, queries)
    quant = net.quantised
    acts = []
    if quant:
        h, dh = _round_interval(x, ex)
    else:
        h, dh = x, np.zeros_like(x)
    acts.append(h)
    gamma = 2.0 * 64 * 2.0**-24
    for layer in range(3):
        w = net.w[layer]
        a = h @ w
        da = dh @ np.abs(w) + gamma * (np.abs(h) @ np.abs(w)) if with_bound else None
        if layer < 2:
            r = np.maximum(a, 0.0)
            if quant:
                if with_bound:
                    lo_, hi_ = np.maximum(a - da, 0.0), np.maximum(a + da, 0.0)
                    h = bf16(r)
                    dh = np.maximum(bf16(hi_) - h, h - bf16(lo_))
                else:
                    h = bf16(r)
            else:
                h = r
                dh = da if with_bound else None
            acts.append(h)
        else:
            out, dout = a, da
    if with_bound:
        return out, acts, dout
    return out, acts


# ---------------------------------------------------------------------------------------------------
# FP8 (OCP e4m3fn) inference: per-tensor power-of-two scales, operands rounded to nearest even after
# saturation at 448, exact products, FP32 accumulation.
# ---------------------------------------------------------------------------------------------------
FP8_MAX = 448.0


def fp8_e4m3(a):
    """Round to OCP FP8 e4m3fn (nearest even) after clamping to +-448; returns float64 values."""
    v = np.clip(np.asarray(a, dtype=np.float64), -FP8_MAX, FP8_MAX)
    m = np.abs(v)
    _, ex = np.frexp(np.where(m > 0, m, 1.0))  # m = f 2^ex, f in [0.5, 1): floor(log2 m) = ex - 1
    e = np.maximum(ex - 1, -6)  # subnormals share the quantum of the smallest normal binade, 2^-9
    q = np.ldexp(1.0, e - 3)
    return np.round(v / q) * q  # half to even


def fp8_pow2_scale(v):
    """Smallest power of two s with v / s <= 448 (v / 448 evaluated in FP32); 1 for v <= 0."""
    v = F32(v)
    if not v > 0:
        return 1.0
    f, e = np.frexp(F32(v / F32(FP8_MAX)))
    return float(np.ldexp(1.0, int(e) - 1 if f == 0.5 else int(e)))


class Fp8Net:
    """The FP8 copy of a guide: FP16 grid as the GPU reads it, weights W / s_w rounded to e4m3 from the
    FP32 master values ([in][out] float64), and the per-tensor weight scales."""

    def __init__(self, params, layout):
        g, w1, w2, w3 = split_params(params, layout)
        self.layout = layout
        self.grid = fp16(g)
        self.sw = [fp8_pow2_scale(np.abs(w).max()) for w in (w1, w2, w3)]
        self.w = [fp8_e4m3(w.astype(np.float64) / s) for w, s in zip((w1, w2, w3), self.sw)]


def fp8_input_scales(amax):
    """Input scales of the three layers from recorded activation maxima: twice the smallest power of two
    that brings the maximum to 448 or below (one power of two of headroom)."""
    return [2.0 * fp8_pow2_scale(a) for a in amax]


def fp8_k_unit(g, j):
    """Input unit of k position 8 g + j in a 32-wide k-block (the GPU's permutation of the units)."""
    return 4 * g + j if j < 4 else 16 + 4 * g + (j - 4)


def fp8_groups(layer):
    """The input units summed by one lane group of one FP8 MFMA (8 products each), per layer."""
    if layer == 0:
        groups = [[fp8_k_unit(g, j) for j in range(8)] for g in range(4)]
        return groups + [list(range(32, 40))]  # normal, wo, 1 (lane group 0 of k-block 1)
    return [[32 * kb + fp8_k_unit(g, j) for j in range(8)] for kb in range(2) for g in range(4)]


# The FP8 MFMA's accumulation, measured on gfx950 with an in-bounds probe: each lane group's 8 products
# are added in a fixed-point adder aligned to the group's largest product, 2^E <= |p| < 2^(E+1), with
# LSB 2^(E - 14); smaller products lose the bits below it (truncation). Group sums and the accumulator
# input are added in FP32. The bound below charges 2^(E - 13) (twice the LSB) per product of a group
# besides its largest, plus FP32 rounding of every group sum.
FP8_GROUP_LSB_BITS = 13


def _group_error(mag_x, w, layer):
    """Bound on the MFMA's deviation from the exact dot product, per (query, output)."""
    err = np.zeros((mag_x.shape[0], w.shape[1]))
    total = np.abs(mag_x) @ np.abs(w)
    for grp in fp8_groups(layer):
        p = np.abs(mag_x[:, grp])[:, :, None] * np.abs(w[grp, :])[None, :, :]  # (n, 8, out)
        m = p.max(axis=1)
        nz = (p > 0).sum(axis=1)
        _, ex = np.frexp(np.where(m > 0, m, 1.0))
        err += np.where(m > 0, np.maximum(nz - 1, 0) * np.ldexp(1.0, ex - 1 - FP8_GROUP_LSB_BITS), 0.0)
    return err + 2.0 * 10 * 2.0**-24 * total


def forward_fp8(net, queries, s_in):
    """FP8 forward: raw outputs (n, 48) and a bound on |GPU - oracle| per output. Every layer input is
    carried as an interval [lo, hi]: the FP32 error of the grid sums, the MFMA's group-aligned
    accumulation (_group_error) and the FP8 rounding of both ends, which is monotone. The nominal value
    rounds the oracle's own exact sums."""
    x, ex = encode(net, queries)
    nom = fp8_e4m3(x / s_in[0])
    lo, hi = fp8_e4m3((x - ex) / s_in[0]), fp8_e4m3((x + ex) / s_in[0])
    for layer in range(3):
        w = net.w[layer]
        wp, wn = np.maximum(w, 0.0), np.minimum(w, 0.0)
        a_nom = nom @ w
        acc_err = _group_error(np.maximum(np.abs(lo), np.abs(hi)), w, layer)
        a_lo = lo @ wp + hi @ wn - acc_err
        a_hi = hi @ wp + lo @ wn + acc_err
        c = s_in[layer] * net.sw[layer]
        if layer < 2:
            nxt = s_in[layer + 1]
            nom = fp8_e4m3(np.maximum(a_nom * c, 0.0) / nxt)
            lo = fp8_e4m3(np.maximum(a_lo * c, 0.0) / nxt)
            hi = fp8_e4m3(np.maximum(a_hi * c, 0.0) / nxt)
        else:
            out = a_nom * c
            bound = np.maximum(a_hi * c - out, out - a_lo * c)
    return out, bound


def head(o):
    """Raw outputs -> (weights (n, 8), mu (n, 8, 3), kappa (n, 8))."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, N_OUT)
    logits = o[:, 0:40:5]
    e = np.exp(logits - logits.max(axis=1, keepdims=True))
    w = e / e.sum(axis=1, keepdims=True)
    raw = np.stack([o[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    nrm = np.linalg.norm(raw, axis=2, keepdims=True)
    mu = np.where(nrm > 1e-12, raw / np.maximum(nrm, 1e-300), np.array([0.0, 0.0, 1.0]))
    kappa = np.clip(np.exp(o[:, 4:40:5]), MIN_KAPPA, MAX_KAPPA)
    return w, mu, kappa


def head_bound(o, do):
    """First-order bounds on |d weights|, |d mu| and |d kappa| for output perturbations up to do, with a
    margin for the FP32 evaluation of the head on the GPU."""
    w, mu, kappa = head(o)
    do = np.asarray(do).reshape(-1, N_OUT)
    dl = do[:, 0:40:5]
    dw = w * (dl + (w * dl).sum(axis=1, keepdims=True)) * 1.5 + 1e-6 * w + 1e-7
    raw = np.stack([o[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    draw = np.stack([do[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    nrm = np.linalg.norm(raw, axis=2)
    dmu = 1.5 * np.linalg.norm(draw, axis=2) / np.maximum(nrm, 1e-30) + 2e-6
    free = (np.exp(o[:, 4:40:5]) > MIN_KAPPA) & (np.exp(o[:, 4:40:5]) < MAX_KAPPA)
    dk = np.where(free, kappa * np.expm1(1.5 * do[:, 4:40:5]), 0.0) + 2e-6 * kappa
    return dw, dmu, dk


def mixture_array(w, mu, kappa):
    """(n, 40) in the C ABI order: w[8], then (mu xyz, kappa) x 8."""
    n = len(w)
    out = np.zeros((n, 40))
    out[:, :8] = w
    lob = np.concatenate([mu, kappa[:, :, None]], axis=2).reshape(n, 32)
    out[:, 8:] = lob
    return out


# ---------------------------------------------------------------------------------------------------
# vMF lobes and mixtures (vmf.h)
# ---------------------------------------------------------------------------------------------------
def vmf_log_norm(kappa):
    """log of kappa / (2 pi (1 - exp(-2 kappa))) (VMF::norm_term, vmf.h:67-73); 1 / (4 pi) at 0."""
    k = np.asarray(kappa, dtype=np.float64)
    small = k < 1e-4
    ks = np.where(small, 1.0, k)
    return np.where(small, -math.log(4 * math.pi), np.log(ks) - math.log(2 * math.pi) - np.log(-np.expm1(-2 * ks)))


def vmf_dlog_norm(kappa):
    k = np.asarray(kappa, dtype=np.float64)
    small = k < 1e-3
    ks = np.where(small, 1.0, k)
    with np.errstate(over="ignore"):
        return np.where(small, 1.0 - k / 3.0, 1.0 / ks - 2.0 / np.expm1(2.0 * ks))


def vmf_eval(mu, kappa, w):
    """VMF::eval (vmf.h:32-39)."""
    d = np.sum(np.asarray(mu) * np.asarray(w), axis=-1)
    return np.exp(vmf_log_norm(kappa) + np.asarray(kappa) * np.minimum(0.0, d - 1.0))


def mixture_log_pdf(mix, dirs):
    mix = np.asarray(mix, dtype=np.float64).reshape(-1, 40)
    dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
    w = mix[:, :8]
    lob = mix[:, 8:].reshape(-1, 8, 4)
    d = np.einsum("nkc,nc->nk", lob[:, :, :3], dirs)
    with np.errstate(divide="ignore"):
        lg = np.where(w > 0, np.log(np.where(w > 0, w, 1.0)) + vmf_log_norm(lob[:, :, 3]) +
                      lob[:, :, 3] * np.minimum(0.0, d - 1.0), -np.inf)
    mx = lg.max(axis=1)
    return mx + np.log(np.exp(lg - mx[:, None]).sum(axis=1))


def mixture_pdf(mix, dirs):
    return np.exp(mixture_log_pdf(mix, dirs))


def orthonormal_basis(n):
    """maths.h:814-821 (Duff et al. 2017), rows of n."""
    n = np.asarray(n, dtype=np.float64)
    sign = np.copysign(1.0, n[..., 2])
    a = -1.0 / (sign + n[..., 2])
    b = n[..., 0] * n[..., 1] * a
    x = np.stack([1.0 + sign * n[..., 0] ** 2 * a, sign * b, -sign * n[..., 0]], axis=-1)
    y = np.stack([b, sign + n[..., 1] ** 2 * a, -n[..., 1]], axis=-1)
    return x, y


def vmf_sample(mu, kappa, u0, u1):
    """VMF::sample (vmf.h:41-56); sample_uniform_sphere (rng.h:284-292) at kappa 0."""
    mu = np.asarray(mu, dtype=np.float64).reshape(-1, 3)
    k = np.asarray(kappa, dtype=np.float64).reshape(-1)
    u0 = np.asarray(u0, dtype=np.float64).reshape(-1)
    u1 = np.asarray(u1, dtype=np.float64).reshape(-1)
    zero = k == 0.0
    ks = np.where(zero, 1.0, k)
    ct = 1.0 + np.log(u0 + np.exp(-2.0 * ks) * (1.0 - u0)) / ks
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = 2 * math.pi * u1
    bx, by = orthonormal_basis(mu)
    w = bx * (np.cos(phi) * st)[:, None] + by * (np.sin(phi) * st)[:, None] + mu * ct[:, None]
    z = 1.0 - 2.0 * u1
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    pu = 2 * math.pi * u0
    uni = np.stack([r * np.cos(pu), r * np.sin(pu), z], axis=1)
    return np.where(zero[:, None], uni, w)


def mixture_select(w, u0):
    """sample_small_distrib (rng.h:437-470) over the lobe weights: (lobe, remapped u0). The remapped
    value is clamped to [0, 1) like the GPU's."""
    w = np.asarray(w, dtype=np.float64).reshape(-1, LOBES)
    u0 = np.asarray(u0, dtype=np.float64).reshape(-1)
    cdf = np.cumsum(w / w.sum(axis=1, keepdims=True), axis=1)
    prev = np.concatenate([np.zeros((len(w), 1)), cdf[:, :-1]], axis=1)
    hit = u0[:, None] < cdf
    any_hit = hit.any(axis=1)
    last = LOBES - 1 - np.argmax((w > 0)[:, ::-1], axis=1)
    sel = np.where(any_hit, np.argmax(hit, axis=1), last)
    rows = np.arange(len(w))
    ur = np.where(any_hit, (u0 - prev[rows, sel]) / (cdf[rows, sel] - prev[rows, sel]), u0)
    return sel, np.clip(ur, 0.0, 1.0 - 2.0**-24)


def mixture_sample(mix, u):
    """sample_vmm (vmf.h:123-127): (directions, lobe index)."""
    mix = np.asarray(mix, dtype=np.float64).reshape(-1, 40)
    u = np.asarray(u, dtype=np.float64).reshape(-1, 2)
    sel, ur = mixture_select(mix[:, :8], u[:, 0])
    lob = mix[:, 8:].reshape(-1, 8, 4)[np.arange(len(mix)), sel]
    return vmf_sample(lob[:, :3], lob[:, 3], ur, u[:, 1]), sel


# ---------------------------------------------------------------------------------------------------
# loss and gradients
# ---------------------------------------------------------------------------------------------------
def head_backward(o, wi, c):
    """(-c log q(wi) per sample, d/d outputs (n, 48)) for the head above."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, N_OUT)
    wi = np.asarray(wi, dtype=np.float64).reshape(-1, 3)
    wi = wi / np.linalg.norm(wi, axis=1, keepdims=True)
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    w, mu, kappa = head(o)
    d = np.einsum("nkc,nc->nk", mu, wi)
    lg = np.log(np.maximum(w, 1e-30)) + vmf_log_norm(kappa) + kappa * np.minimum(0.0, d - 1.0)
    mx = lg.max(axis=1, keepdims=True)
    logq = (mx + np.log(np.exp(lg - mx).sum(axis=1, keepdims=True)))[:, 0]
    r = np.exp(lg - logq[:, None])
    g = np.zeros_like(o)
    g[:, 0:40:5] = -c[:, None] * (r - w)
    e = np.exp(o[:, 4:40:5])
    free = (e > MIN_KAPPA) & (e < MAX_KAPPA)
    dk = -c[:, None] * r * (vmf_dlog_norm(kappa) + np.minimum(0.0, d - 1.0))
    g[:, 4:40:5] = np.where(free, dk * kappa, 0.0)
    raw = np.stack([o[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    nrm = np.linalg.norm(raw, axis=2)
    gm = (wi[:, None, :] - mu * d[:, :, None]) * (-c[:, None] * r * kappa / np.maximum(nrm, 1e-30))[:, :, None]
    gm = np.where((nrm > 1e-12)[:, :, None], gm, 0.0)
    for k in range(LOBES):
        g[:, 5 * k + 1:5 * k + 4] = gm[:, k]
    active = c > 0
    return np.where(active, -c * logq, 0.0), np.where(active[:, None], g, 0.0)


def loss(params, layout, queries, wi, c, quantised=False):
    net = Net(params, layout, quantised)
    o, _ = forward(net, queries)
    l, _ = head_backward(o, wi, c)
    return float(l.sum())


def gradient(params, layout, queries, wi, c, quantised=False):
    """(loss, gradient over all parameters) by back-propagation. quantised=True follows the GPU: BF16
    output gradients and hidden-layer gradients, FP32 weight gradients and grid scatter."""
    net = Net(params, layout, quantised)
    q = np.asarray(queries, dtype=F32).reshape(-1, 9)
    o, acts = forward(net, q)
    x, h1, h2 = acts
    lval, g = head_backward(o, wi, c)
    rq = bf16 if quantised else (lambda a: a)
    w1, w2, w3 = net.w
    g = rq(g)
    dw3 = h2.T @ g
    dh2 = rq(np.where(h2 > 0, g @ w3.T, 0.0))
    dw2 = h1.T @ dh2
    dh1 = rq(np.where(h1 > 0, dh2 @ w2.T, 0.0))
    dw1 = x.T @ dh1
    dx = dh1 @ w1.T
    grad = np.zeros(layout.n_grid + MLP_SIZE)
    gg = grad[: layout.n_grid].reshape(-1, FEATURES)
    for l, (idx, w) in enumerate(layout.corners(q[:, 0:3])):
        for f in range(FEATURES):
            np.add.at(gg[:, f], idx.ravel(), (w * dx[:, 2 * l + f][:, None]).ravel())
    o0 = layout.n_grid
    grad[o0:o0 + W1_SIZE] = dw1.ravel()
    grad[o0 + W1_SIZE:o0 + W1_SIZE + W2_SIZE] = dw2.ravel()
    grad[o0 + W1_SIZE + W2_SIZE:] = dw3.ravel()
    return float(lval.sum()), grad


# ---------------------------------------------------------------------------------------------------
# Adam (gpu/shaders/adam.slang:39-50)
# ---------------------------------------------------------------------------------------------------
def adam_step(params, grad, m1, m2, it, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8):
    """One step at iteration `it` (starting at 1): returns (params, m1, m2), FP32 arithmetic."""
    p, g, m, v = (np.asarray(a, dtype=F32) for a in (params, grad, m1, m2))
    b1, b2, lr_, ep = F32(beta1), F32(beta2), F32(lr), F32(eps)
    mt = g + b1 * (m - g)
    vt = g * g + b2 * (v - g * g)
    bc1 = F32(1) - F32(np.float32(beta1) ** F32(it))
    bc2 = F32(1) - F32(np.float32(beta2) ** F32(it))
    p = p - lr_ * (mt / bc1) / (np.sqrt(vt / bc2) + ep)
    return p.astype(F32), mt.astype(F32), vt.astype(F32)
