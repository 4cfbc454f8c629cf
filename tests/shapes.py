"""Small axis-aligned test meshes built in numpy (triangle soups, 9 floats per triangle)."""
import numpy as np


def box_faces(half, n=1, skip=(), inward=False, centre=(0.0, 0.0, 0.0)):
    """Surface of the cube [-half, half]^3 + centre, each face split into n x n quads of two triangles, wound
    outward (inward=True: reversed).  skip: faces left out, named '+x', '-x', '+y', '-y', '+z', '-z'."""
    tris = []
    g = np.linspace(-half, half, n + 1)
    for axis in range(3):
        for sgn in (1.0, -1.0):
            name = ("+" if sgn > 0 else "-") + "xyz"[axis]
            if name in skip:
                continue
            u, v = (axis + 1) % 3, (axis + 2) % 3
            for i in range(n):
                for j in range(n):
                    quad = []
                    for a, b in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = sgn * half
                        p[u], p[v] = g[a], g[b]
                        quad.append(p)
                    t0, t1 = [quad[0], quad[1], quad[2]], [quad[0], quad[2], quad[3]]
                    if sgn < 0:  # (u, v, axis) is right handed: reverse on the negative faces
                        t0, t1 = t0[::-1], t1[::-1]
                    tris += [t0, t1]
    xyz = np.array(tris, np.float64) + np.asarray(centre, np.float64)
    if inward:
        xyz = xyz[:, ::-1, :]
    return np.ascontiguousarray(xyz, np.float32)


def nested_boxes(count, first, step, alternate=True):
    """`count` concentric cubes of half sizes first + k * step (2 triangles per face), the odd ones wound
    inward when `alternate`, so that the winding number alternates 1, 0, 1, ... from the outside in."""
    return np.concatenate([box_faces(first + k * step, 1, inward=alternate and (count - 1 - k) % 2 == 1)
                           for k in range(count)])
