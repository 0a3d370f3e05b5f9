"""Synthetic workloads of the benchmark configurations, as dense Batch buffers.

Values are produced on the host with numpy (bit-exact and deterministic from the
seed) and handed to the oracle and to the device path unchanged.

plain_checker   one 32^3 tree, v = ((x+y+z)&1)+1 at all voxels
terrain         64x64 grid of 32^3 trees sharing one interner; 2-D gradient-noise
                height field (seeded 256-entry permutation, 4 octaves, lacunarity 2,
                gain 0.5, base period 128 voxels); h = 8 + floor(16 n); id 1 below
                the surface, 2 at it, nothing above
uneven          trees of depth drawn from a set; dense ((x+y+z) mod 65535)+1 fills
                or sparse uniform random edits at a drawn density
"""
import numpy as np

from .morton import encode_path

_PERM_CACHE = {}


def morton_order(depth):
    """perm such that dense_yzx.ravel()[perm[m]] is the voxel with Morton path m."""
    if depth in _PERM_CACHE:
        return _PERM_CACHE[depth]
    n = 1 << depth
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m = encode_path(xyz)
    perm = np.empty(n * n * n, dtype=np.int64)
    perm[m] = np.arange(n * n * n)
    _PERM_CACHE[depth] = perm
    return perm


def dense_to_batch(dense_yzx, depth):
    """Dense [N,N,N] (y,z,x) voxel values -> (masks [P,2] u8, values [P,8]) as a Batch
    with just_set(v) for every non-default voxel (no clears recorded)."""
    flat = np.asarray(dense_yzx).reshape(-1)[morton_order(depth)]
    values = flat.reshape(-1, 8)
    bits = (values != 0).astype(np.uint8)
    set_mask = np.packbits(bits, axis=1, bitorder="little")[:, 0]
    masks = np.zeros((values.shape[0], 2), dtype=np.uint8)
    masks[:, 0] = set_mask
    return masks, np.ascontiguousarray(values)


def dense_many_to_batches(dense, depth):
    """[T, N, N, N] -> masks [T, P, 2], values [T, P, 8] (vectorized over trees)."""
    t = dense.shape[0]
    flat = dense.reshape(t, -1)[:, morton_order(depth)]
    values = np.ascontiguousarray(flat.reshape(t, -1, 8))
    bits = (values != 0).astype(np.uint8)
    masks = np.zeros(values.shape[:2] + (2,), dtype=np.uint8)
    masks[:, :, 0] = np.packbits(bits, axis=2, bitorder="little")[:, :, 0]
    return masks, values


def plain_checker(depth=5):
    n = 1 << depth
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    dense = (((x + y + z) & 1) + 1).astype(np.uint8)
    return dense


# ---------------------------------------------------------------- terrain

def _perm_table(seed):
    rng = np.random.default_rng(seed)
    p = rng.permutation(256).astype(np.int64)
    return np.concatenate([p, p])


_GRADS = np.array([[1, 1], [-1, 1], [1, -1], [-1, -1], [1, 0], [-1, 0], [0, 1], [0, -1]],
                  dtype=np.float64)


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def gradient_noise_2d(X, Z, perm):
    """Classic 2-D gradient (Perlin) noise at float coordinates, roughly in [-1, 1]."""
    xi = np.floor(X).astype(np.int64)
    zi = np.floor(Z).astype(np.int64)
    xf = X - xi
    zf = Z - zi
    xi &= 255
    zi &= 255

    def grad(ix, iz, dx, dz):
        h = perm[perm[ix] + iz] & 7
        g = _GRADS[h]
        return g[..., 0] * dx + g[..., 1] * dz

    n00 = grad(xi, zi, xf, zf)
    n10 = grad(xi + 1, zi, xf - 1.0, zf)
    n01 = grad(xi, zi + 1, xf, zf - 1.0)
    n11 = grad(xi + 1, zi + 1, xf - 1.0, zf - 1.0)
    u = _fade(xf)
    w = _fade(zf)
    a = n00 + u * (n10 - n00)
    b = n01 + u * (n11 - n01)
    return a + w * (b - a)


def terrain_heights(grid=64, chunk=32, seed=1, octaves=4, base_period=128.0, lacunarity=2.0,
                    gain=0.5):
    """u8 column heights [grid*chunk, grid*chunk] indexed [Z, X] in world voxels."""
    perm = _perm_table(seed)
    size = grid * chunk
    Z, X = np.meshgrid(np.arange(size, dtype=np.float64), np.arange(size, dtype=np.float64),
                       indexing="ij")
    total = np.zeros_like(X)
    amp, period, norm = 1.0, base_period, 0.0
    for _ in range(octaves):
        total += amp * gradient_noise_2d(X / period + 0.5, Z / period + 0.5, perm)
        norm += amp
        amp *= gain
        period /= lacunarity
    n = (total / norm + 1.0) * 0.5
    n = np.clip(n, 0.0, np.nextafter(1.0, 0.0))
    return (8 + np.floor(16.0 * n)).astype(np.uint8)


def terrain_dense(heights, cx, cz, chunk=32, below=1, surface=2):
    """Dense (y,z,x) u8 voxels of tree (cx, cz)."""
    h = heights[cz * chunk:(cz + 1) * chunk, cx * chunk:(cx + 1) * chunk].astype(np.int32)  # [z,x]
    y = np.arange(chunk, dtype=np.int32)[:, None, None]
    hh = h[None, :, :]
    out = np.where(y < hh, below, np.where(y == hh, surface, 0)).astype(np.uint8)
    return out


def terrain_batches(grid=64, chunk=32, seed=1, below=1, surface=2, trees=None):
    """masks [T,P,2], values [T,P,8], edits per tree for trees (cz*grid+cx order)."""
    depth = int(np.log2(chunk))
    heights = terrain_heights(grid, chunk, seed)
    if trees is None:
        trees = range(grid * grid)
    trees = list(trees)
    dense = np.empty((len(trees), chunk, chunk, chunk), dtype=np.uint8)
    for k, t in enumerate(trees):
        cz, cx = divmod(t, grid)
        dense[k] = terrain_dense(heights, cx, cz, chunk, below, surface)
    masks, values = dense_many_to_batches(dense, depth)
    edits = (dense != 0).reshape(len(trees), -1).sum(axis=1)
    return masks, values, edits, dense


# ---------------------------------------------------------------- uneven

def uneven_spec(n_trees, depths=(5, 6), p_dense=0.5, densities=(0.001, 0.01, 0.1), seed=3):
    """Per tree: (depth, kind, density, seed) drawn from the seed."""
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(n_trees):
        d = int(rng.choice(depths))
        dense = bool(rng.random() < p_dense)
        dens = float(rng.choice(densities))
        spec.append((d, "dense" if dense else "sparse", dens, int(rng.integers(1 << 31))))
    return spec


def uneven_tree_edits(depth, kind, density, seed):
    """(xyz int32 [n,3], values u16 [n]) in submission order for one tree."""
    n = 1 << depth
    if kind == "dense":
        y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
        vals = ((xyz.sum(axis=1) % 65535) + 1).astype(np.uint16)
        return xyz, vals
    rng = np.random.default_rng(seed)
    k = max(1, int(round(density * n * n * n)))
    xyz = rng.integers(0, n, size=(k, 3)).astype(np.int32)
    vals = rng.integers(1, 65536, size=k).astype(np.uint16)  # 0 excluded (a clear is a no-op)
    return xyz, vals


def uneven_dense(depth):
    """Dense (y,z,x) u16 voxels of a dense tree: ((x+y+z) mod 65535) + 1."""
    n = 1 << depth
    y, z, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (((x + y + z) % 65535) + 1).astype(np.uint16)


def uneven_batches(spec):
    """Per tree (masks [P,2] u8, values [P,8] u16, edits) for an uneven_spec list."""
    from .morton import apply_edits_to_dense, dense_batch_arrays
    out = []
    for depth, kind, density, seed in spec:
        if kind == "dense":
            m, v = dense_to_batch(uneven_dense(depth), depth)
            out.append((m, v, 1 << (3 * depth)))
        else:
            xyz, vals = uneven_tree_edits(depth, kind, density, seed)
            m, v = dense_batch_arrays(depth, np.uint16)
            apply_edits_to_dense(m, v, xyz, vals)
            out.append((m, v, len(vals)))
    return out
