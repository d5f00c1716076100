"""TEST INFRASTRUCTURE — FP64 numpy restatement of the path-guiding network.

The guiding proposal has no counterpart in the reference (it extends the
plain BSDF+NEE sampling of src/kernels/pathtrace.cu:548-905); this file is the
CPU oracle its HIP kernels are checked against. It defines:

  * the 80-wide input: a 12-frequency sin/cos encoding of the scene-normalised
    position (72), the shading normal (3), the outgoing direction (3), the
    roughness (1) and the normalised wavelength (1);
  * an MLP 80 -> 64 -> 64 -> 64 -> 64 -> 41 with ReLU on the hidden layers;
  * the output as 8 von Mises-Fisher lobes, 5 numbers each: a mean direction
    (normalised), log sharpness (clamped) and a mixture logit (softmax), and
    a selection logit c: the renderer draws from the lobes with probability
    a = SEL_MIN + (SEL_MAX - SEL_MIN) sigmoid(c), otherwise from the BSDF;
    in the fixed-selection mode the selection output is held at zero (its
    weights and bias are zero and get no gradient), so a = 0.5 and the
    network is the 40-output mixture network;
  * the loss of one training record, weight * -log(a q(d) + (1 - a) p_bsdf(d)),
    whose batch sum divided by the weight sum is minimised: a KL divergence
    between the density of the sampled contribution (BSDF x cosine x incident
    radiance, the record weight being that product over the sampling pdf) and
    the one-sample MIS density the renderer actually samples;
  * its analytic gradient and Adam.

Record layout (16 floats): position[0:3] in [0,1]^3, normal[3:6], wo[6:9],
roughness[9], wavelength in [0,1][10], sampled direction[11:14], weight[14],
BSDF-sampler density of the direction[15].
"""
import numpy as np

N_FREQ = 12
N_IN = 80
N_IN_PAD = 96          # K padded to a multiple of 32 for the MFMA tiles
WIDTH = 64
N_LOBES = 8
N_MIX = 5 * N_LOBES    # 40 lobe outputs
N_OUT = N_MIX + 1      # and the selection logit
SEL_MIN, SEL_MAX = 0.1, 0.9
N_OUT_PAD = 48
LAYERS = [(N_IN, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, N_OUT)]
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -3.0, 7.0
REC = 16


def n_params():
    return sum(i * o + o for i, o in LAYERS)


def encode(rec):
    """(n, 16) records -> (n, 80) features."""
    rec = np.asarray(rec, np.float64)
    n = len(rec)
    f = np.zeros((n, N_IN))
    col = 0
    for a in range(3):
        for k in range(N_FREQ):
            x = rec[:, a] * (2.0 ** k)
            # sin(pi 2^k p) and cos(pi 2^k p): the argument reduced modulo 2
            r = x - 2.0 * np.floor(x * 0.5)
            f[:, col] = np.sin(np.pi * r)
            f[:, col + 1] = np.cos(np.pi * r)
            col += 2
    f[:, 72:75] = rec[:, 3:6]
    f[:, 75:78] = rec[:, 6:9]
    f[:, 78] = rec[:, 9]
    f[:, 79] = rec[:, 10]
    return f


def init_params(seed=1):
    """He-uniform weights, zero biases; returns a list of (W[out, in], b[out])."""
    rng = np.random.default_rng(seed)
    ps = []
    for fi, fo in LAYERS:
        lim = np.sqrt(6.0 / fi)
        ps.append((rng.uniform(-lim, lim, (fo, fi)), np.zeros(fo)))
    return ps


def flatten(ps):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in ps])


def unflatten(v):
    v = np.asarray(v, np.float64)
    ps, o = [], 0
    for fi, fo in LAYERS:
        W = v[o:o + fi * fo].reshape(fo, fi)
        o += fi * fo
        b = v[o:o + fo]
        o += fo
        ps.append((W, b))
    return ps


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(ps, x, emulate_bf16=False):
    """Returns the raw 41 outputs and the per-layer inputs (for backward).
    With emulate_bf16 the layer inputs and weights are rounded to bfloat16 as
    the matrix-core kernels store them (products accumulate exactly here)."""
    q = bf16 if emulate_bf16 else (lambda a: a)
    h = q(x)
    acts = [h]
    for li, (W, b) in enumerate(ps):
        z = h @ q(W).T + b
        if li < len(ps) - 1:
            h = q(np.maximum(z, 0.0))
        else:
            h = z
        acts.append(h)
    return h, acts


def mixture(out):
    """Raw outputs -> (mu (n,8,3), kappa (n,8), weights (n,8), pieces for backward)."""
    o = np.asarray(out)[:, :N_MIX].reshape(-1, N_LOBES, 5)
    raw = o[..., 0:3]
    nrm = np.maximum(np.linalg.norm(raw, axis=-1, keepdims=True), 1e-6)
    mu = raw / nrm
    lk = np.clip(o[..., 3], LOG_KAPPA_MIN, LOG_KAPPA_MAX)
    kappa = np.exp(lk)
    lg = o[..., 4]
    lg = lg - lg.max(axis=1, keepdims=True)
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w, nrm


def selection(out):
    """Probability of drawing from the lobes rather than the BSDF."""
    c = np.asarray(out, np.float64)[:, N_MIX]
    return SEL_MIN + (SEL_MAX - SEL_MIN) / (1.0 + np.exp(-c))


def log_norm(kappa):
    """log of kappa / (2 pi (1 - exp(-2 kappa))), the vMF normalisation."""
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2.0 * kappa))


def log_pdf(out, d):
    mu, kappa, w, _ = mixture(out)
    lc = np.log(w) + log_norm(kappa) + kappa * (np.einsum("nkc,nc->nk", mu, d) - 1.0)
    m = lc.max(axis=1, keepdims=True)
    return (m + np.log(np.exp(lc - m).sum(axis=1, keepdims=True)))[:, 0]


def pdf(out, d):
    return np.exp(log_pdf(out, d))


def sample(out, u):
    """u: (n, 3) uniforms -> directions drawn from the mixture (lobe by u0)."""
    mu, kappa, w, _ = mixture(out)
    c = np.cumsum(w, axis=1)
    j = np.minimum((u[:, :1] > c).sum(axis=1), N_LOBES - 1)
    idx = np.arange(len(u))
    m, k = mu[idx, j], kappa[idx, j]
    cz = 1.0 + np.log(u[:, 1] + (1.0 - u[:, 1]) * np.exp(-2.0 * k)) / k
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    ph = 2 * np.pi * u[:, 2]
    t, bt = frame(m)
    return (t * (sz * np.cos(ph))[:, None] + bt * (sz * np.sin(ph))[:, None] + m * cz[:, None])


def frame(n):
    s = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1 + s * n[:, 0] ** 2 * a, s * b, -s * n[:, 0]], 1)
    bt = np.stack([b, s + n[:, 1] ** 2 * a, -n[:, 1]], 1)
    return t, bt


def fix_selection(ps):
    """Parameters of the fixed-selection mode: the last layer's selection row
    (weights and bias of output N_MIX) set to zero."""
    ps = [(W.copy(), b.copy()) for W, b in ps]
    ps[-1][0][N_MIX, :] = 0.0
    ps[-1][1][N_MIX] = 0.0
    return ps


def loss_and_dout(out, rec, fixed_selection=False):
    """Batch loss sum(w * -log(a q + (1 - a) p_bsdf)) / sum(w) and its
    gradient w.r.t. the raw outputs. With fixed_selection, a = 0.5 and the
    selection output gets no gradient."""
    rec = np.asarray(rec, np.float64)
    d = rec[:, 11:14]
    wt = rec[:, 14]
    pb = rec[:, 15]
    W = max(wt.sum(), 1e-30)
    o = np.asarray(out)[:, :N_MIX].reshape(-1, N_LOBES, 5)
    mu, kappa, w, nrm = mixture(out)
    dot = np.einsum("nkc,nc->nk", mu, d)
    lc = np.log(w) + log_norm(kappa) + kappa * (dot - 1.0)
    m = lc.max(axis=1, keepdims=True)
    lse = (m + np.log(np.exp(lc - m).sum(axis=1, keepdims=True)))[:, 0]
    q = np.exp(lse)
    sig = 1.0 / (1.0 + np.exp(-np.asarray(out, np.float64)[:, N_MIX]))
    a = np.full_like(sig, 0.5) if fixed_selection else SEL_MIN + (SEL_MAX - SEL_MIN) * sig
    P = np.maximum(a * q + (1.0 - a) * pb, 1e-30)
    loss = float(np.sum(wt * -np.log(P)) / W)
    g = np.exp(lc - lse[:, None])                      # responsibilities within the lobes
    s = (wt / W * a * q / P)[:, None]                  # share of the lobes in P
    dlogit = -s * (g - w)
    # d(-log q)/d kappa = -g (1/kappa - 2/(e^{2k}-1) + dot - 1); kappa = exp(lk)
    with np.errstate(over="ignore"):               # expm1 -> inf for large kappa: 2 / inf = 0, the limit
        dlnorm = 1.0 / kappa - 2.0 / np.expm1(2.0 * kappa)
    dk = -s * g * (dlnorm + dot - 1.0)
    lk_raw = o[..., 3]
    inside = (lk_raw > LOG_KAPPA_MIN) & (lk_raw < LOG_KAPPA_MAX)
    dlk = dk * kappa * inside
    # mu = raw / |raw|: d/draw = (I - mu mu^T) / |raw| * d/dmu, d/dmu = -g kappa d
    dmu = -(s * g * kappa)[..., None] * d[:, None, :]
    proj = dmu - mu * np.sum(dmu * mu, axis=-1, keepdims=True)
    draw = proj / nrm
    dout = np.zeros((len(rec), N_OUT))
    dl = dout[:, :N_MIX].reshape(-1, N_LOBES, 5)
    dl[..., 0:3] = draw
    dl[..., 3] = dlk
    dl[..., 4] = dlogit
    dout[:, :N_MIX] = dl.reshape(-1, N_MIX)
    # selection: d(-log P)/da = -(q - p_bsdf) / P, da/dc = (SEL_MAX - SEL_MIN) sig (1 - sig)
    if not fixed_selection:
        dout[:, N_MIX] = -(wt / W) * (q - pb) / P * (SEL_MAX - SEL_MIN) * sig * (1.0 - sig)
    return loss, dout


def backward(ps, acts, dout, emulate_bf16=False):
    q = bf16 if emulate_bf16 else (lambda a: a)
    grads = [None] * len(ps)
    dh = q(dout)
    for li in range(len(ps) - 1, -1, -1):
        W, b = ps[li]
        x = acts[li]
        grads[li] = (dh.T @ x, dh.sum(axis=0))
        if li > 0:
            dh = q((dh @ q(W)) * (acts[li] > 0))
    return grads


def loss_grad(ps, rec, emulate_bf16=False, fixed_selection=False):
    if fixed_selection:
        ps = fix_selection(ps)
    x = encode(rec)
    out, acts = forward(ps, x, emulate_bf16)
    loss, dout = loss_and_dout(out, rec, fixed_selection)
    return loss, backward(ps, acts, dout, emulate_bf16)


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)
