"""Full-size checks of whole-layer convolution results (test infrastructure).

The FP64 oracle's direct loops are too slow for a whole benchmark layer, but
every entry of a weight gradient is one dot product over all the layer's
output pixels, and every image of a forward / backward-input result depends
on its own image only.  So a full-size result is checked by:

  * ``dw_entries``: sampled weight-gradient entries, each an FP64 dot product
    over all ``B * Ho * Wo`` pixels of the shifted input channel and the output
    gradient channel (``owl_ndarray_conv_impl.h:675-888``: ``dw[a, c, q, o] =
    sum_{b, i, j} x[b, i sH + a - pT, j sW + c - pL, q] dy[b, i, j, o]``, zero
    outside the image), computed on the host, independently of the oracle's
    loops and of the GPU;
  * ``pick_images``: images away from the front of the batch, whose forward /
    backward-input results go through the oracle image by image.
"""
import numpy as np

from owl_geometry import lead_pad


def sample_dw_index(ksp, ci, co, n, seed):
    """n distinct (a, c, q, o) entries of a [kh, kw, Ci, Co] weight gradient."""
    rng = np.random.default_rng(seed)
    total = ksp[0] * ksp[1] * ci * co
    flat = rng.choice(total, size=min(n, total), replace=False)
    return np.stack(np.unravel_index(flat, (ksp[0], ksp[1], ci, co)), axis=1)


def dw_entries(x, dy, ksp, stride, idx):
    """FP64 weight-gradient entries ``idx`` (rows (a, c, q, o)) of the 2-D
    convolution of x [B, H, W, Ci] with a [kh, kw, Ci, Co] kernel at the given
    stride (rate 1), SAME or VALID as the shapes imply (the leading pad the C
    side derives from them)."""
    B, H, W, _ = x.shape
    _, Ho, Wo, _ = dy.shape
    kh, kw = ksp
    sh, sw = stride
    ph, pw = lead_pad(sh, Ho, kh, H), lead_pad(sw, Wo, kw, W)
    Hp, Wp = max(H + ph, sh * (Ho - 1) + kh), max(W + pw, sw * (Wo - 1) + kw)
    out = np.empty(len(idx))
    xcache, gcache = {}, {}
    for n, (a, c, q, o) in enumerate(idx):
        if q not in xcache:
            xc = np.zeros((B, Hp, Wp))
            xc[:, ph:ph + H, pw:pw + W] = x[..., q]
            xcache[q] = xc
        if o not in gcache:
            gcache[o] = np.ascontiguousarray(dy[..., o], dtype=np.float64)
        win = xcache[q][:, a:a + sh * (Ho - 1) + 1:sh, c:c + sw * (Wo - 1) + 1:sw]
        out[n] = float(np.tensordot(win, gcache[o], axes=3))
    return out


def sampled_rel(got, ref):
    """rel-L2 of the sampled entries; 1.0 for an all-zero result."""
    got = np.asarray(got, np.float64)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def pick_images(batch, n, seed, skip=2):
    """n distinct image indices from [skip, batch)."""
    rng = np.random.default_rng(seed)
    return sorted(int(i) for i in rng.choice(np.arange(skip, batch), size=n, replace=False))
