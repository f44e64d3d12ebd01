"""Synthetic inputs for the synthesis workloads (no datasets are available).

SinNoise(seed, stream, W, H, C): per channel, four sinusoids with amplitude
a ~ U(0.5, 1), frequencies fx, fy ~ +-U(0.5, 8) cycles per image and phase
U(0, 2pi); v = 0.9 * (1/2 + 1/2 * sum(a sin(.)) / sum(a)) + 0.1 * U(0, 1),
so every channel, alpha included, lies in [0, 1].

Orientation guides are unit vectors at angle 2pi * SmoothNoise (a cubic
upsampled 8x8 lattice of U(0,1)), delivered as the structure tensor
(c^2, c s, s^2) the reference's orientation input expects
(G/nlInputElements.cpp:33).  Distance transforms are the distance to 64
uniform seed points divided by the maximum.
"""
import numpy as np
from scipy import ndimage


def _rng(seed, stream):
    return np.random.default_rng([int(seed), int(stream), 0x9A1D])


def sin_noise(seed, stream, w, h, c=4):
    rng = _rng(seed, stream)
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    out = np.empty((h, w, c), np.float32)
    for k in range(c):
        a = rng.uniform(0.5, 1.0, 4)
        fx = rng.uniform(0.5, 8.0, 4) * rng.choice([-1.0, 1.0], 4)
        fy = rng.uniform(0.5, 8.0, 4) * rng.choice([-1.0, 1.0], 4)
        ph = rng.uniform(0.0, 2 * np.pi, 4)
        s = np.zeros((h, w))
        for i in range(4):
            # sin(u + v) = sin u cos v + cos u sin v with u along x, v along y
            u = 2 * np.pi * fx[i] * x / w + ph[i]
            v = 2 * np.pi * fy[i] * y / h
            s += a[i] * (np.outer(np.cos(v), np.sin(u)) + np.outer(np.sin(v), np.cos(u)))
        v = 0.9 * (0.5 + 0.5 * s / a.sum()) + 0.1 * rng.uniform(0.0, 1.0, (h, w))
        out[..., k] = v
    return out


def smooth_noise(seed, stream, w, h, lattice=8):
    rng = _rng(seed, stream)
    g = rng.uniform(0.0, 1.0, (lattice, lattice))
    z = ndimage.zoom(g, (h / lattice, w / lattice), order=3, mode="grid-wrap", grid_mode=True)
    return z[:h, :w]


def orientation_tensor(seed, stream, w, h):
    th = 2 * np.pi * smooth_noise(seed, stream, w, h)
    c, s = np.cos(th), np.sin(th)
    return np.stack([c * c, c * s, s * s], axis=-1).astype(np.float32)


def object_ids(seed, stream, w, h, regions=4):
    """Object-ID image (int32): the Voronoi cell of `regions` uniform seed
    points, cell k carrying the ID 1009 * k + 7 (IDs are arbitrary integers,
    as the reference's r + 255 g + 255^2 b codes are)."""
    rng = _rng(seed, stream)
    px = rng.uniform(0, w, regions)
    py = rng.uniform(0, h, regions)
    from scipy.spatial import cKDTree
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    _, k = cKDTree(np.stack([px, py], axis=1)).query(np.stack([xx.ravel(), yy.ravel()], axis=1))
    return (1009 * k.reshape(h, w) + 7).astype(np.int32)


def distance_transform(seed, stream, w, h, points=64):
    rng = _rng(seed, stream)
    px = rng.uniform(0, w, points)
    py = rng.uniform(0, h, points)
    from scipy.spatial import cKDTree
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    d, _ = cKDTree(np.stack([px, py], axis=1)).query(np.stack([xx.ravel(), yy.ravel()], axis=1))
    d = d.reshape(h, w)
    return (d / d.max()).astype(np.float32)


def culling_dist_trans(seed, stream, w, h, lo=30.0, hi=70.0):
    """A distance transform in pixel units, lo + (hi - lo) * SmoothNoise, for
    inputs that reach the culling rules: with the cull distance 50 an image
    pixel is culled above 50 (isImagePixelCulled), a previous-frame pixel
    invalidates advection above 42 (cull - 8) and an exemplar texel rejects
    candidates above 62.5 (1.25 * cull); the default range straddles all
    three."""
    return (lo + (hi - lo) * smooth_noise(seed, stream, w, h)).astype(np.float32)


def style_pair(w, h, seed, guides=False):
    """Exemplar input/output pair (streams 0, 1) and optional guides."""
    st = {"base": sin_noise(seed, 0, w, h), "output": sin_noise(seed, 1, w, h)}
    if guides:
        st["orientation"] = orientation_tensor(seed, 10, w, h)
        st["dist_trans"] = distance_transform(seed, 11, w, h)
    return st


def still_frame(w, h, seed, guides=False):
    """A target frame (stream 2) with optional orientation and DT guides."""
    fr = {"color": sin_noise(seed, 2, w, h)}
    if guides:
        fr["orientation"] = orientation_tensor(seed, 20, w, h)
        fr["dist_trans"] = distance_transform(seed, 21, w, h)
    return fr


def _motion(seed, t, w, h):
    """M_t(x) = R(0.5 deg)(x - c) + c - x + (3, 1) + 2 * SmoothNoise2(t)."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    a = np.deg2rad(0.5)
    rx = np.cos(a) * (xx - cx) - np.sin(a) * (yy - cy) + cx - xx
    ry = np.sin(a) * (xx - cx) + np.cos(a) * (yy - cy) + cy - yy
    nx = 2.0 * (smooth_noise(seed, 100 + 2 * t, w, h) - 0.5)
    ny = 2.0 * (smooth_noise(seed, 101 + 2 * t, w, h) - 0.5)
    return np.stack([rx + 3.0 + nx, ry + 1.0 + ny], axis=-1)


def animation(w, h, frames, seed, guides=True):
    """Frames whose colour follows the motion field; vel_f[t] = M_t and
    vel_b[t](x) = -M_{t-1}(x - M_{t-1}(x)) (zero for the first frame)."""
    base = sin_noise(seed, 2, w + 64, h + 64)
    out = []
    shift = np.zeros(2)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    dt = distance_transform(seed, 21, w, h) if guides else None
    for t in range(frames):
        m = _motion(seed, t, w, h)
        sx, sy = shift
        col = np.empty((h, w, 4), np.float32)
        for k in range(4):
            col[..., k] = ndimage.map_coordinates(base[..., k], [yy + 32 - sy, xx + 32 - sx], order=1, mode="wrap")
        fr = {"color": col, "vel_f": m.astype(np.float32)}
        if t > 0:
            mp = _motion(seed, t - 1, w, h)
            qx = xx - mp[..., 0]
            qy = yy - mp[..., 1]
            bx = -ndimage.map_coordinates(mp[..., 0], [qy, qx], order=1, mode="nearest")
            by = -ndimage.map_coordinates(mp[..., 1], [qy, qx], order=1, mode="nearest")
            fr["vel_b"] = np.stack([bx, by], axis=-1).astype(np.float32)
        else:
            fr["vel_b"] = np.zeros((h, w, 2), np.float32)
        if guides:
            fr["orientation"] = orientation_tensor(seed, 20 + t, w, h)
            fr["dist_trans"] = dt
        out.append(fr)
        shift += (3.0, 1.0)
    return out


def _log_uniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def batch_specs(n=64, seed=4, iterations=6):
    """Jobs of the batched workload: area log-uniform 128^2..1024^2, aspect
    0.5..2, patch radius 2 or 3 (5x5 / 7x7), 3 to 5 levels, exemplar area
    log-uniform 256^2..1024^2 (square); job j uses seed `seed + j`."""
    rng = np.random.default_rng([int(seed), 0xBA7C])
    specs = []
    for j in range(n):
        area = _log_uniform(rng, 128 * 128, 1024 * 1024)
        aspect = _log_uniform(rng, 0.5, 2.0)
        w = max(16, int(round(np.sqrt(area * aspect))))
        h = max(16, int(round(np.sqrt(area / aspect))))
        r = int(rng.choice([2, 3]))
        levels = int(rng.integers(3, 6))
        s = int(round(np.sqrt(_log_uniform(rng, 256 * 256, 1024 * 1024))))
        levels = min(levels, int(np.floor(np.log2(min(w, h)))), int(np.floor(np.log2(s))))
        specs.append(dict(w=w, h=h, r=r, levels=levels, sw=s, sh=s, seed=seed + j, iterations=iterations))
    return specs


def batch_job(spec):
    """(frame, style) of one batched job, with orientation and DT guides."""
    return (still_frame(spec["w"], spec["h"], seed=spec["seed"], guides=True),
            style_pair(spec["sw"], spec["sh"], seed=spec["seed"], guides=True))


def serving_specs(n=32, seed=5):
    """Sequences of the serving workload: frame sizes uniform 256..1920 x
    256..1080, square exemplars 512..1024; sequence k uses seed `seed + k`."""
    rng = np.random.default_rng([int(seed), 0x5E7F])
    specs = []
    for k in range(n):
        specs.append(dict(w=int(rng.integers(256, 1921)), h=int(rng.integers(256, 1081)),
                          s=int(rng.integers(512, 1025)), seed=seed + k))
    return specs


def serving_sequence(spec, frames):
    """(frames, style) of one served sequence: the animation motion model
    (rotation 0.5 deg/frame, translation (3, 1), smooth noise 2 px)."""
    return (animation(spec["w"], spec["h"], frames, seed=spec["seed"]),
            style_pair(spec["s"], spec["s"], seed=spec["seed"], guides=True))
