"""ORACLE -- test infrastructure only: the CPU restatement of the compositing's
face pruning (pf_prune_faces, powerfoam_amd/csrc/power.hip) that tests/ use as
the checker; the product never imports it. No reference counterpart: the
reference walks every face of a site's list (rasterize.py:1049-1071).

`prune` follows the kernel's FP32 arithmetic operation by operation (numpy
float32 scalars: IEEE single, correctly rounded division and square root, no
fused multiply-add), so its outputs equal the kernel's bit for bit.
`interval` is the FP64 interval of a ray inside a site's ball and cell (the
geometry the compositing computes, rasterize.py:1028-1086 without the texel
surface); the tests use it to show that the faces `prune` drops change no
interval.
"""

import numpy as np

F = np.float32


def _half(u16):
    return F(np.array(u16, np.uint16).view(np.float16))


def face(spheres, adjacency, diff, s, e):
    """(nx, ny, nz, w) of edge e of site s in float32, as the kernel forms it:
    n the neighbour delta, w = ((|n|^2 + r^2) - r_a^2) / 2."""
    c = spheres[s]
    if diff is not None:
        raw = diff[e].view(np.uint16) if diff.dtype == np.float16 else diff[e]
        nx, ny, nz, ra = (_half(raw[k]) for k in range(4))
    else:
        a = spheres[adjacency[e]]
        nx, ny, nz, ra = F(a[0] - c[0]), F(a[1] - c[1]), F(a[2] - c[2]), F(a[3])
    n2 = F(F(F(nx * nx) + F(ny * ny)) + F(nz * nz))
    w = F(F(0.5) * F(F(n2 + F(c[3] * c[3])) - F(ra * ra)))
    return nx, ny, nz, w


def dropped(spheres, adjacency, diff, offsets, s, margin):
    """Per edge of site s: True when pf_prune_faces leaves it out."""
    a0, a1 = int(offsets[s]), int(offsets[s + 1])
    c = spheres[s].astype(np.float32)
    m = F(margin)
    r2 = F(c[3] * c[3])
    lim = F(F(F(1.0) + m) * F(F(1.0) + m))
    mr = F(m * c[3])
    fs = [face(spheres, adjacency, diff, s, e) for e in range(a0, a1)]
    out = []
    for i, (fx, fy, fz, fw) in enumerate(fs):
        drop = False
        if fw > F(0.0):
            nf2 = F(F(F(fx * fx) + F(fy * fy)) + F(fz * fz))
            if F(fw * fw) >= F(F(r2 * lim) * nf2):
                drop = True
            else:
                q = F(fw / nf2)
                rho = F(np.sqrt(max(F(r2 - F(fw * q)), F(0.0))))
                for k, (gx, gy, gz, gw) in enumerate(fs):
                    if k == i or not (gw > F(0.0)):
                        continue
                    ng2 = F(F(F(gx * gx) + F(gy * gy)) + F(gz * gz))
                    d = F(F(F(gx * fx) + F(gy * fy)) + F(gz * fz))
                    perp2 = max(F(ng2 - F(d * F(d / nf2))), F(0.0))
                    lhs = F(F(F(d * q) - F(rho * F(np.sqrt(perp2)))) - gw)
                    if lhs > F(mr * F(np.sqrt(ng2))):
                        drop = True
                        break
        out.append(drop)
    return np.array(out, bool)


def prune(spheres, adjacency, offsets, diff=None, margin=1e-3):
    """(adjacency_out, diff_out or None, count): each site's kept faces first
    in list order, then the dropped ones from the end of its range backwards,
    as the kernel writes them."""
    spheres = np.asarray(spheres, np.float32)
    n = spheres.shape[0]
    adj_out = np.array(adjacency, np.int32, copy=True)
    diff_out = None if diff is None else np.array(diff, copy=True)
    count = np.zeros(n, np.int32)
    for s in range(n):
        a0, a1 = int(offsets[s]), int(offsets[s + 1])
        if a1 == a0:
            continue
        d = dropped(spheres, adjacency, diff, offsets, s, margin)
        kept = np.nonzero(~d)[0] + a0
        gone = np.nonzero(d)[0] + a0
        order = np.concatenate([kept, gone[::-1]])
        adj_out[a0:a1] = np.asarray(adjacency)[order]
        if diff is not None:
            diff_out[a0:a1] = np.asarray(diff)[order]
        count[s] = kept.size
    return adj_out, diff_out, count


def interval(o, d, c, r, faces):
    """FP64 [t_near, t_far] of the ray o + t d (|d| = 1) inside the ball
    (c, r) and the cell n.(x - c) < w of every (n, w) in faces; None if
    empty."""
    oc = o - c
    b = float(oc @ d)
    h2 = r * r - float(oc @ oc - b * b)
    if h2 < 0.0:
        return None
    h = np.sqrt(h2)
    tn, tf = max(-b - h, 0.0), -b + h
    if tf < 0.0:
        return None
    for n, w in faces:
        dp = float(n @ d)
        num = w - float(n @ oc)
        if dp == 0.0:
            if num <= 0.0:
                return None
            continue
        t = num / dp
        if dp > 0.0:
            tf = min(tf, t)
        else:
            tn = max(tn, t)
    return (tn, tf) if tn < tf else None
