"""Synthetic inputs of the benchmark workloads (host side, numpy).

water_box: the graph-partitioned SP2 workload -- atoms at uniform random
positions in a periodic cube at liquid-water number density, a few orbitals
per atom, hopping -exp(-r / decay) between orbitals of atoms closer than the
cutoff.  The choices the workload statement leaves open are fixed here:

* number density 0.1003 atoms / A^3 (all three atoms of 0.0334 molecules / A^3),
  so 10,000 atoms fill a cube of side 46.4 A;
* orbital a of atom i couples to orbital b of atom j (i != j, r_ij < cutoff)
  with -exp(-r_ij / decay) * s_a * s_b, s = (1.0, 0.7, 0.4);
* on-site energies (-0.5, 0.0, 0.5) eV per orbital plus U[-0.1, 0.1] noise,
  no intra-atom coupling;
* orbital index = 3 * atom + a, atoms in generation order.

Neighbours are found with a cell list (cells of side >= cutoff) under the
minimum-image convention.
"""
from __future__ import annotations

import numpy as np

WATER_ATOM_DENSITY = 0.1003  # atoms / A^3


def water_box(natoms=10000, density=WATER_ATOM_DENSITY, cutoff=4.0, decay=1.5,
              scales=(1.0, 0.7, 0.4), onsite=(-0.5, 0.0, 0.5), seed=0):
    """Returns (rows, cols, vals, norb, box): the upper and lower triangle and
    the diagonal of the Hamiltonian as COO arrays (rows/cols int64)."""
    rng = np.random.default_rng(seed)
    box = (natoms / density) ** (1.0 / 3.0)
    pos = rng.uniform(0.0, box, (natoms, 3))
    nb = 3
    norb = nb * natoms
    ncell = max(1, int(box // cutoff))
    cell = np.floor(pos / (box / ncell)).astype(np.int64) % ncell
    cid = (cell[:, 0] * ncell + cell[:, 1]) * ncell + cell[:, 2]
    order = np.argsort(cid, kind="stable")
    starts = np.searchsorted(cid[order], np.arange(ncell ** 3 + 1))
    ii, jj, rr = [], [], []
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    for cx in range(ncell):
        for cy in range(ncell):
            for cz in range(ncell):
                me = order[starts[(cx * ncell + cy) * ncell + cz]:
                           starts[(cx * ncell + cy) * ncell + cz + 1]]
                if me.size == 0:
                    continue
                seen = set()
                for dx, dy, dz in offs:
                    key = (((cx + dx) % ncell) * ncell + (cy + dy) % ncell) * ncell + (cz + dz) % ncell
                    if key in seen:
                        continue
                    seen.add(key)
                    other = order[starts[key]:starts[key + 1]]
                    if other.size == 0:
                        continue
                    d = pos[me][:, None, :] - pos[other][None, :, :]
                    d -= box * np.rint(d / box)
                    r = np.sqrt(np.sum(d * d, axis=-1))
                    a, b = np.nonzero((r < cutoff) & (me[:, None] < other[None, :]))
                    ii.append(me[a])
                    jj.append(other[b])
                    rr.append(r[a, b])
    ii = np.concatenate(ii) if ii else np.zeros(0, np.int64)
    jj = np.concatenate(jj) if jj else np.zeros(0, np.int64)
    rr = np.concatenate(rr) if rr else np.zeros(0)
    s = np.asarray(scales, dtype=float)
    t = -np.exp(-rr / decay)
    rows, cols, vals = [], [], []
    for a in range(nb):
        for b in range(nb):
            rows.append(nb * ii + a)
            cols.append(nb * jj + b)
            vals.append(t * s[a] * s[b])
    r_up = np.concatenate(rows)
    c_up = np.concatenate(cols)
    v_up = np.concatenate(vals)
    diag = np.tile(np.asarray(onsite, dtype=float), natoms) + rng.uniform(-0.1, 0.1, norb)
    d_idx = np.arange(norb, dtype=np.int64)
    rows = np.concatenate([r_up, c_up, d_idx])
    cols = np.concatenate([c_up, r_up, d_idx])
    vals = np.concatenate([v_up, v_up, diag])
    return rows, cols, vals, norb, box


def dense(rows, cols, vals, n, out=None):
    """Scatter COO arrays into a dense n x n float64 array (numpy or torch)."""
    if out is None:
        out = np.zeros((n, n))
    out[rows, cols] = vals
    return out
