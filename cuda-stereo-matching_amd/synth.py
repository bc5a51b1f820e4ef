"""Seeded synthetic stereo pairs of known disparity.

Recipe (identical on every machine with the same numpy: PCG64 streams and
plain float64 arithmetic only):
  1. left image: N(0,1) white noise, separable Gaussian blur (sigma 1 px,
     radius 3, reflected borders), standardised to mean 128 / std 40,
     clipped to [0, 255];
  2. ground truth: `regions` Voronoi cells around uniform seeds, each a
     slanted plane d = a*u + b*v + c with c ~ U[4, dmax-8] and
     a, b ~ U[-slope, slope], clipped to [0, dmax-1];
  3. right image: the left forward-warped to x - floor(d + 0.5) with a z-buffer
     that keeps the larger disparity (then the larger source x); holes get
     fresh blurred noise;
  4. optional occluded strips: vertical strips of the left image (width
     4..16) covering `occlusion` of the columns, whose right-image
     correspondents are replaced by fresh noise;
  5. N(0, 2^2) noise on both images, rounded and clipped to uint8.
"""
import numpy as np

_G = np.exp(-0.5 * np.arange(-3, 4, dtype=np.float64) ** 2)
GAUSS7 = _G / _G.sum()


def _blur(x):
    """Separable 7-tap Gaussian (sigma 1) with reflected borders."""
    p = np.pad(x, 3, mode="reflect")
    h = sum(GAUSS7[k] * p[:, k:k + x.shape[1]] for k in range(7))
    return sum(GAUSS7[k] * h[k:k + x.shape[0], :] for k in range(7))


def _texture(rng, H, W):
    n = _blur(rng.standard_normal((H + 6, W + 6)))[3:3 + H, 3:3 + W]
    n = (n - n.mean()) / n.std() * 40.0 + 128.0
    return np.clip(n, 0.0, 255.0)


def ground_truth(rng, W, H, dmax, regions, slope):
    sx = rng.uniform(0, W, regions)
    sy = rng.uniform(0, H, regions)
    a = rng.uniform(-slope, slope, regions)
    b = rng.uniform(-slope, slope, regions)
    c = rng.uniform(4, dmax - 8, regions)
    u = np.arange(W, dtype=np.float64)[None, :]
    v = np.arange(H, dtype=np.float64)[:, None]
    best = np.full((H, W), np.inf)
    label = np.zeros((H, W), dtype=np.int32)
    for k in range(regions):
        dist = (u - sx[k]) ** 2 + (v - sy[k]) ** 2
        m = dist < best
        best[m] = dist[m]
        label[m] = k
    d = a[label] * u + b[label] * v + c[label]
    return np.clip(d, 0.0, dmax - 1.0)


def make_pair(W, H, dmax, regions=8, seed=0, slope=0.05, occlusion=0.0, noise=2.0):
    """Returns (left u8 [H,W], right u8 [H,W], ground-truth disparity f64 [H,W])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    left = _texture(rng, H, W)
    gt = ground_truth(rng, W, H, dmax, regions, slope)
    di = np.floor(gt + 0.5).astype(np.int64)
    ys, xs = np.mgrid[0:H, 0:W]
    xr = xs - di
    ok = xr >= 0
    key = np.full(H * W, -1, dtype=np.int64)
    tgt = (ys * W + xr)[ok]
    np.maximum.at(key, tgt, (di * (W + 1) + xs)[ok])
    fresh = _texture(rng, H, W)
    right = fresh.copy().reshape(-1)
    hit = key >= 0
    src_x = key[hit] % (W + 1)
    src_y = np.nonzero(hit)[0] // W
    right[hit] = left[src_y, src_x]
    right = right.reshape(H, W)
    if occlusion > 0:
        fresh2 = _texture(rng, H, W)
        covered = 0
        while covered < occlusion * W:
            w = int(rng.integers(4, 17))
            x0 = int(rng.integers(0, max(1, W - w)))
            cols = np.arange(x0, min(W, x0 + w))
            covered += len(cols)
            yy, xx = np.meshgrid(np.arange(H), cols, indexing="ij")
            xt = xx - di[yy, xx]
            m = xt >= 0
            right[yy[m], xt[m]] = fresh2[yy[m], xt[m]]
    l8 = np.clip(np.rint(left + rng.normal(0.0, noise, (H, W))), 0, 255).astype(np.uint8)
    r8 = np.clip(np.rint(right + rng.normal(0.0, noise, (H, W))), 0, 255).astype(np.uint8)
    return l8, r8, gt


# Named workloads: (width, height, disp_max, regions, slope, occlusion, seed)
CONFIGS = {
    "450x375_d64": dict(W=450, H=375, dmax=63, regions=8, slope=0.05, occlusion=0.0, seed=0),
    "1242x375_d128": dict(W=1242, H=375, dmax=127, regions=16, slope=0.05, occlusion=0.05, seed=1),
    "2816x1880_d256": dict(W=2816, H=1880, dmax=255, regions=32, slope=0.1, occlusion=0.0, seed=2),
}

# Batch mix: sizes drawn 3:2:1 from these shapes, the same generator.
BATCH_SHAPES = [(640, 480, 63), (1242, 375, 127), (1920, 1080, 255)]


def config_pair(name):
    c = CONFIGS[name]
    l, r, gt = make_pair(c["W"], c["H"], c["dmax"], c["regions"], c["seed"], c["slope"], c["occlusion"])
    return l, r, gt, c["dmax"]


def batch_shapes(n, seed=3000):
    """Shapes of an n-pair batch: pair i draws from BATCH_SHAPES with p = 1/2, 1/3, 1/6."""
    out = []
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        k = int(rng.choice(3, p=[3 / 6, 2 / 6, 1 / 6]))
        out.append(BATCH_SHAPES[k])
    return out


def batch_pairs(n, seed=3000):
    """n pairs of the batch mix: (left, right, disp_max) per pair, pair i seeded seed+i."""
    pairs = []
    for i, (W, H, dmax) in enumerate(batch_shapes(n, seed)):
        l, r, _ = make_pair(W, H, dmax, regions=8, seed=seed + i, slope=0.05)
        pairs.append((l, r, dmax))
    return pairs
