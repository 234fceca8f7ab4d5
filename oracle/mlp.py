"""FP64 numpy restatement of the learned band-CG preconditioner.

TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU leg): the checker for the
device MLP kernels, never part of the product path.

There is no reference code for this part (it is new in this build); the
specification it restates is the one the device kernels implement:

  features of plane wave g in band b (E_b = band kinetic energy
  sum_g kin_g |psi_bg|^2 / sum_g |psi_bg|^2, Vm / Vs = mean / standard
  deviation of the local potential on the grid):
      f = [kin_g / E_b, log(E_b), Vm, Vs]
  network 4 -> 64 -> 64 -> 64 -> 1, SiLU activations, weights and biases
  rounded to bfloat16 (round to nearest even) for the forward pass, all
  arithmetic in the accumulation type;
  preconditioner scale s = t(x) exp(o) / E_b with x = kin_g / E_b and t the
  Teter-Payne-Allan polynomial (uspp_preconditioner.rs:148-180), so a
  network whose output is 0 is the per-band TPA preconditioner;
  loss over residual / correction pairs (r, e), band by band (band_len
  consecutive pairs per band; bands whose mask w <= 0 are unused), with
  a = -Re(conj(r) e), c = |r|^2:
      L = (1/n_bands) sum_{used b} [1 - (sum s a)^2 / (sum s^2 c sum |e|^2)]
  and its gradient with respect to the FP32 master parameters (straight
  through the bfloat16 rounding).

Parameter vector layout (8705 floats): W1[64][4], b1[64], W2[64][64],
b2[64], W3[64][64], b3[64], w4[64], b4.
"""
import numpy as np

H = 64
N_IN = 4
SIZES = [("W1", (H, N_IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
         ("W3", (H, H)), ("b3", (H,)), ("w4", (H,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def to_bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    rounding = ((a >> 16) & 1) + 0x7FFF
    b = ((a + rounding) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def lcg_uniform(seed, n):
    """Uniform [0, 1) from the 48-bit LCG used for every seeded quantity."""
    s = int(seed) & 0xFFFFFFFFFFFF
    out = np.empty(n)
    for i in range(n):
        s = (s * 25214903917 + 11) & 0xFFFFFFFFFFFF
        out[i] = s / 281474976710656.0
    return out


def init_params(seed=11):
    """Uniform Glorot init of W1..W3 from the LCG; biases, w4 and b4 zero
    (the initial network is the per-band TPA preconditioner)."""
    d = {}
    u = lcg_uniform(seed, H * N_IN + 2 * H * H)
    o = 0
    for name, (fo, fi) in (("W1", (H, N_IN)), ("W2", (H, H)), ("W3", (H, H))):
        lim = np.sqrt(6.0 / (fo + fi))
        d[name] = ((2.0 * u[o:o + fo * fi] - 1.0) * lim).reshape(fo, fi)
        o += fo * fi
    for name in ("b1", "b2", "b3", "w4"):
        d[name] = np.zeros(H)
    d["b4"] = np.zeros(1)
    return np.asarray(pack(d), np.float32)


def tpa(x):
    x2 = x * x
    x3 = x2 * x
    num = 27.0 + 18.0 * x + 12.0 * x2 + 8.0 * x3
    return num / (num + 16.0 * x2 * x2)


def features(kin, ekin_b, vmean, vstd):
    """(n_b, n_pw, 4) features from per-PW kinetic energies and per-band E_b."""
    kin = np.asarray(kin, np.float64)
    e = np.asarray(ekin_b, np.float64)[:, None]
    f = np.empty((e.shape[0], kin.shape[0], 4))
    f[..., 0] = kin[None, :] / e
    f[..., 1] = np.log(e)
    f[..., 2] = vmean
    f[..., 3] = vstd
    return f


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return sg * (1.0 + z * (1.0 - sg))


def forward(theta, f):
    """Network output o for feature rows f (N, 4), bf16 weights, FP64 math."""
    p = unpack(to_bf16(theta))
    z1 = f @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    z3 = h2 @ p["W3"].T + p["b3"]
    h3 = silu(z3)
    o = h3 @ p["w4"] + p["b4"][0]
    return o, (p, z1, h1, z2, h2, z3, h3)


def scale(theta, f):
    """Preconditioner scale s = t(x) exp(o) / E_b for feature rows f."""
    o, _ = forward(theta, f)
    x = f[:, 0]
    eb = np.exp(f[:, 1])
    return tpa(x) * np.exp(o) / eb


def band_loss(s, r, e, w, band_len):
    """L and dL/ds of the band-mean misalignment loss for scales s."""
    nb = len(s) // band_len
    a = -np.real(np.conj(r) * e)
    c = np.abs(r) ** 2
    A = (s * a).reshape(nb, band_len).sum(1)
    B = (s * s * c).reshape(nb, band_len).sum(1)
    E = (np.abs(e) ** 2).reshape(nb, band_len).sum(1)
    used = (w.reshape(nb, band_len)[:, 0] > 0) & (B > 0) & (E > 0)
    Bs = np.where(used, B, 1.0)
    Es = np.where(used, E, 1.0)
    L = float(np.sum(np.where(used, 1.0 - A * A / (Bs * Es), 0.0)) / nb)
    g1 = np.where(used, 2.0 * A / (Bs * Es), 0.0) / nb
    g2 = np.where(used, A / Bs, 0.0)
    dLds = np.repeat(g1, band_len) * (np.repeat(g2, band_len) * s * c - a)
    return L, dLds


def loss_grad(theta, f, r, e, w, band_len):
    """Band loss L and dL/dtheta (FP64) for pairs (r, e)."""
    o, (p, z1, h1, z2, h2, z3, h3) = forward(theta, f)
    s = tpa(f[:, 0]) * np.exp(o) / np.exp(f[:, 1])
    L, dLds = band_loss(s, r, e, w, band_len)
    do = dLds * s                                       # ds/do = s
    g = {"b4": np.array([do.sum()]), "w4": h3.T @ do}
    dz3 = np.outer(do, p["w4"]) * dsilu(z3)
    g["W3"] = dz3.T @ h2
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ p["W3"]) * dsilu(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * dsilu(z1)
    g["W1"] = dz1.T @ f
    g["b1"] = dz1.sum(0)
    return L, pack(g)


def adam_step(theta, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t counts from 1) of FP32 master parameters."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
