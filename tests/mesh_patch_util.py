"""A seamed test mesh for mesh-based patches: an n x n vertex grid in the
y = 0 plane whose middle column is duplicated (two copies at the same
positions, as an OBJ file splits vertices along a texture seam), with
per-corner material coordinates and the two top corners coloured red."""
import numpy as np


def seamed_grid(n=9, side=1.0, seam=None):
    seam = n // 2 if seam is None else seam
    xs = np.linspace(-side / 2, side / 2, n, dtype=np.float32)
    verts, uvs, ids = [], [], {}
    for j in range(n):          # columns along x
        for i in range(n):      # rows along z
            copies = 2 if j == seam else 1
            for c in range(copies):
                ids[(j, i, c)] = len(verts)
                verts.append((xs[j], 0.0, xs[i]))
                uvs.append((xs[j] + (0.0 if c == 0 else 0.0), xs[i]))
    tris = []
    for j in range(n - 1):
        for i in range(n - 1):
            # left column of a seam cell uses copy 1 of the seam column on its right side
            a = ids[(j, i, 1 if j == seam else 0)]
            b = ids[(j + 1, i, 0)]
            c = ids[(j + 1, i + 1, 0)]
            d = ids[(j, i + 1, 1 if j == seam else 0)]
            if (i + j) % 2 == 0:
                tris += [(a, b, c), (a, c, d)]
            else:
                tris += [(a, b, d), (b, c, d)]
    colors = np.zeros((len(verts), 3), dtype=np.float32)
    colors[ids[(0, n - 1, 0)], 0] = 1.0
    colors[ids[(n - 1, n - 1, 0)], 0] = 1.0
    return (np.array(verts, dtype=np.float32), np.array(tris, dtype=np.uint32), np.array(uvs, dtype=np.float32),
            colors)
