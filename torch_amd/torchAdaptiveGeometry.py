"""Balanced dyadic leaf meshes for the adaptive scheme (host side).

Mirrors python/cfdSolver/torchAdaptiveGeometry.py: the options and their
checks (``settings``, :7-41), child keys and the canonical leaf order
(:44-58), exact box volumes and face areas of every geometry (:61-112) and
``LeafMesh`` (:115-372): tree validation, the face list (one shared face per
subface, boundary faces, periodic seams), face areas and centre offsets,
the 2:1 balance check, the least-squares gradient matrices and the bounded
least-squares gradients (used on the host only when a regrid prolongs
data; the per-step gradients run on the device).

The mesh is host data: the device receives its face list, per-cell face
incidence and metric arrays (torch_amd.torchAdaptive) and keeps the state.
"""

from itertools import product

import numpy as np

DEFAULTS = dict(max_level=2, min_level=0, refine_threshold=0.08, coarsen_threshold=0.02, regrid_every=3,
                max_cells=100000, order=2)


def settings(**overrides):
    """Adaptivity options with the reference's checks."""
    out = dict(DEFAULTS)
    if set(overrides) - set(out):
        raise ValueError("Unknown adaptivity option")
    out.update(overrides)
    for key in ("max_level", "min_level", "regrid_every", "max_cells", "order"):
        v = out[key]
        if not isinstance(v, (int, np.integer)) or isinstance(v, bool):
            raise ValueError("AMR counts and levels must be integers")
        out[key] = int(v)
    if not (0 <= out["min_level"] <= out["max_level"] <= 8 and out["regrid_every"] >= 1
            and 2 <= out["max_cells"] <= 2_000_000 and out["order"] in (1, 2)):
        raise ValueError("Invalid AMR levels, interval, capacity or order")
    low, high = out["coarsen_threshold"], out["refine_threshold"]
    if not np.isfinite([low, high]).all() or not 0 <= low < high <= 1:
        raise ValueError("Require 0 <= coarsen threshold < refine <= 1")
    return out


def children(key):
    """Child keys of a leaf key (level, i[, j[, k]]), x bit fastest."""
    level, *index = key
    return [(level + 1, *(2 * i + b for i, b in zip(index, bits[::-1]))) for bits in product((0, 1), repeat=len(index))]


def canonical_order(keys):
    """Permutation sorting leaves by level, then by the last index down to the first."""
    keys = np.asarray(keys, dtype=np.int64)
    return np.lexsort(tuple(keys[:, k] for k in range(1, keys.shape[1])) + (keys[:, 0],))


def canonical(keys):
    """Leaves in canonical order."""
    keys = np.asarray(keys, dtype=np.int64)
    return keys[canonical_order(keys)]


def volume(c, low, high):
    d, g = c["dimension"], c.get("geometry", "cartesian")
    width = high - low
    if g == "cartesian":
        return np.prod(width, axis=1)
    sq = (high[:, 0] ** 2 - low[:, 0] ** 2) / 2
    if g == "spherical":
        cube = (high[:, 0] ** 3 - low[:, 0] ** 3) / 3
        if d == 1:
            return 4 * np.pi * cube
        out = cube * (np.cos(low[:, 1]) - np.cos(high[:, 1]))
        return out * (width[:, 2] if d == 3 else 2 * np.pi)
    if d == 1:
        return 2 * np.pi * sq
    if g == "cylindrical" and d == 2:
        return 2 * np.pi * sq * width[:, 1]
    return sq * np.prod(width[:, 1:], axis=1)


def area(c, direction, position, low, high):
    d, g = c["dimension"], c.get("geometry", "cartesian")
    width = high - low
    if g == "cartesian":
        return np.prod(np.delete(width, direction, axis=1), axis=1)
    sq = (high[:, 0] ** 2 - low[:, 0] ** 2) / 2
    if d == 1:
        return 4 * np.pi * position ** 2 if g == "spherical" else 2 * np.pi * position
    if g == "spherical":
        phi = width[:, 2] if d == 3 else 2 * np.pi
        if direction == 0:
            return position ** 2 * (np.cos(low[:, 1]) - np.cos(high[:, 1])) * phi
        if direction == 1:
            return sq * np.sin(position) * phi
        return sq * width[:, 1]
    if g == "cylindrical" and d == 2:
        return 2 * np.pi * position * width[:, 1] if direction == 0 else 2 * np.pi * sq
    if direction == 0:
        return position * np.prod(width[:, 1:], axis=1)
    if direction == 1:
        return width[:, 0] * (width[:, 2] if d == 3 else 1)
    return sq * width[:, 1]


def _roots(counts):
    d = len(counts)
    idx = np.indices(tuple(counts[::-1])).reshape(d, -1)[::-1].T
    return np.column_stack((np.zeros(len(idx), dtype=np.int64), idx))


class LeafMesh:
    """A forest of quad/octrees with one shared flux per subface."""

    def __init__(self, config, keys=None, validate=True):
        self.config = config
        self.dimension = d = config["dimension"]
        self.options = settings(**config["amr"])
        self.max_level = self.options["max_level"]
        counts = np.array(config["counts_xyz"], dtype=np.int64)
        if np.prod(counts) > self.options["max_cells"]:
            raise ValueError("AMR capacity is smaller than the base mesh")
        if np.prod(counts) * 2 ** (d * self.options["min_level"]) > self.options["max_cells"]:
            raise ValueError("AMR capacity cannot cover the minimum level")
        keys = _roots(counts) if keys is None else np.asarray(keys)
        if (keys.ndim != 2 or keys.shape[1] != d + 1 or keys.dtype.kind not in "iu" or len(keys) == 0
                or len(keys) > self.options["max_cells"]):
            raise ValueError("Invalid adaptive leaf keys/count")
        self.keys = canonical(keys)
        self.levels = self.keys[:, 0]
        if np.any((self.levels < 0) | (self.levels > self.max_level)):
            raise ValueError("Leaf level is outside configured bounds")
        if np.any(self.keys[:, 1:] < 0) or np.any(self.keys[:, 1:] >= counts * (2 ** self.levels[:, None])):
            raise ValueError("Leaf indices are outside the domain")
        self.counts = counts
        sizes = [int(np.prod(counts)) * 2 ** (d * level) for level in range(self.max_level + 1)]
        self._offsets = np.concatenate(([0], np.cumsum(sizes)[:-1])).astype(np.int64)
        codes = np.empty(len(self.keys), dtype=np.int64)
        for level in np.unique(self.levels):
            at = self.levels == level
            codes[at] = self._codes(level, self.keys[at, 1:])
        self._code_order = np.argsort(codes, kind="stable")
        self._sorted_codes = codes[self._code_order]
        if np.any(self._sorted_codes[1:] == self._sorted_codes[:-1]):
            raise ValueError("Duplicate leaf keys")
        if validate:
            self._check_tree(counts)
        scale = 2 ** (self.max_level - self.levels)
        self.integer_low = self.keys[:, 1:] * scale[:, None]
        self.integer_high = self.integer_low + scale[:, None]
        self.total = counts * 2 ** self.max_level
        self.unit = np.array(config["lengths_xyz"]) / self.total
        self.origin = np.array(config["origins_xyz"])
        self.low = self.origin + self.integer_low * self.unit
        self.high = self.origin + self.integer_high * self.unit
        self.centers = 0.5 * (self.low + self.high)
        self.widths = self.high - self.low
        self.volumes = volume(config, self.low, self.high)
        if not np.isfinite(self.volumes).all() or np.any(self.volumes <= 0):
            raise ValueError("Nonpositive adaptive volumes")
        self.divergence = np.stack([(area(config, k, self.high[:, k], self.low, self.high)
                                     - area(config, k, self.low[:, k], self.low, self.high)) / self.volumes
                                    for k in range(d)])
        self._topology()

    def _check_tree(self, counts):
        """Every root is covered exactly once: complete sibling groups merge
        upward level by level (a merged parent that is itself a leaf means
        overlapping cells), and what remains must be the roots."""
        d = self.dimension
        levels, idx = self.levels.copy(), self.keys[:, 1:].copy()
        for level in range(self.max_level, 0, -1):
            at = np.flatnonzero(levels == level)
            if not len(at):
                continue
            parents = idx[at] // 2
            codes = self._codes(level - 1, parents)
            uniq, first, inverse, count = np.unique(codes, return_index=True, return_inverse=True,
                                                   return_counts=True)
            full = count == 2 ** d
            present = levels == level - 1
            if present.any() and np.isin(uniq[full], self._codes(level - 1, idx[present])).any():
                raise ValueError("Overlapping adaptive cells")
            keep = np.ones(len(levels), dtype=bool)
            keep[at[full[inverse.ravel()]]] = False
            levels = np.concatenate((levels[keep], np.full(int(full.sum()), level - 1)))
            idx = np.concatenate((idx[keep], parents[first[full]]))
        if not (np.all(levels == 0) and len(levels) == int(np.prod(counts))):
            raise ValueError("Adaptive leaves overlap or leave domain holes")

    def _codes(self, level, idx):
        """Integer codes of the keys (level, idx...), idx (n, d) in x, y, z order."""
        dims = self.counts * 2 ** level
        return self._offsets[level] + np.ravel_multi_index(tuple(idx.T[::-1]), tuple(dims[::-1]))

    def _find(self, points):
        """Leaf containing each point (n, d), in finest-level integer units;
        the coarsest level holding a point wins (the reference's _find order)."""
        found = np.full(len(points), -1, dtype=np.int64)
        for level in range(self.max_level + 1):
            pending = np.flatnonzero(found < 0)
            if not len(pending):
                break
            idx = np.floor(points[pending] / 2 ** (self.max_level - level)).astype(np.int64)
            codes = self._codes(level, idx)
            pos = np.minimum(np.searchsorted(self._sorted_codes, codes), len(self._sorted_codes) - 1)
            hit = self._sorted_codes[pos] == codes
            found[pending[hit]] = self._code_order[pos[hit]]
        if np.any(found < 0):
            raise ValueError("Adaptive mesh has an uncovered neighbor")
        return found

    def _face_list(self):
        """The reference's face order (torchAdaptiveGeometry.py:197-240),
        built with array operations: per leaf in canonical order and per
        direction, the high face (a domain-boundary face, the shared face
        with the neighbour, or its subfaces in depth-first order where the
        neighbour is finer), then the low domain-boundary face."""
        c, d = self.config, self.dimension
        periodic = [c["face_boundaries"][2 * k]["type"] == "periodic" for k in range(d)]
        width = max(2 ** (d - 1), 1)
        depth_cap = self.max_level + 1
        parts = []

        def emit(owner, k, part, left, right, lo, hi, position, side, code):
            parts.append(dict(owner=owner, k=np.full(len(owner), k), part=np.full(len(owner), part), left=left,
                              right=right, lo=lo, hi=hi, position=position, side=np.full(len(owner), side),
                              code=code))

        for k in range(d):
            others = np.arange(d) != k
            transverse = [a for a in range(d) if a != k]
            position = self.integer_high[:, k]
            edge = (position == self.total[k]) & (not periodic[k])
            ib = np.flatnonzero(edge)
            emit(ib, k, 0, ib, np.full(len(ib), -1), self.integer_low[ib], self.integer_high[ib], position[ib], 1,
                 np.zeros(len(ib), dtype=np.int64))
            owner = np.flatnonzero(~edge)
            lo, hi, pos = self.integer_low[owner].copy(), self.integer_high[owner].copy(), position[owner]
            code = np.zeros(len(owner), dtype=np.int64)
            for depth in range(depth_cap + 1):
                if not len(owner):
                    break
                if depth == depth_cap:
                    raise ValueError("Adaptive mesh has an uncovered neighbor")
                point = 0.5 * (lo + hi)
                point[:, k] = (pos + 0.5) % self.total[k]
                j = self._find(point)
                covers = (np.all(self.integer_low[j][:, others] <= lo[:, others], axis=1)
                          & np.all(self.integer_high[j][:, others] >= hi[:, others], axis=1))
                emit(owner[covers], k, 0, owner[covers], j[covers], lo[covers], hi[covers], pos[covers], 0,
                     code[covers] * width ** (depth_cap - depth))
                split = ~covers
                owner, lo, hi, pos, code = owner[split], lo[split], hi[split], pos[split], code[split]
                middle = (lo + hi) // 2
                subs = []
                for b, bits in enumerate(product((0, 1), repeat=len(transverse))):
                    sub_lo, sub_hi = lo.copy(), hi.copy()
                    for a, bit in zip(transverse, bits):
                        if bit:
                            sub_lo[:, a] = middle[:, a]
                        else:
                            sub_hi[:, a] = middle[:, a]
                    subs.append((owner, sub_lo, sub_hi, pos, code * width + b))
                owner, lo, hi, pos, code = (np.concatenate([s[f] for s in subs]) for f in range(5))
            il = np.flatnonzero((self.integer_low[:, k] == 0) & (not periodic[k]))
            emit(il, k, 1, np.full(len(il), -1), il, self.integer_low[il], self.integer_high[il],
                 np.zeros(len(il), dtype=np.int64), -1, np.zeros(len(il), dtype=np.int64))
        cat = {f: np.concatenate([p[f] for p in parts]) for f in parts[0]}
        order = np.lexsort((cat["code"], cat["part"], cat["k"], cat["owner"]))
        return {f: v[order] for f, v in cat.items()}

    def _topology(self):
        c, d = self.config, self.dimension
        faces = self._face_list()
        self.left, self.right = faces["left"].astype(np.int64), faces["right"].astype(np.int64)
        self.directions, self.sides = faces["k"].astype(np.int64), faces["side"].astype(np.int64)
        lows, highs, positions = faces["lo"], faces["hi"], faces["position"]
        self.face_low = self.origin + np.array(lows) * self.unit
        self.face_high = self.origin + np.array(highs) * self.unit
        self.face_positions = self.origin[self.directions] + np.array(positions) * self.unit[self.directions]
        self.face_areas = np.empty(len(self.left))
        center = 0.5 * (self.face_low + self.face_high)
        center[np.arange(len(self.left)), self.directions] = self.face_positions
        self.offset_left = center - self.centers[np.maximum(self.left, 0)]
        self.offset_right = center - self.centers[np.maximum(self.right, 0)]
        for k in range(d):
            mask = self.directions == k
            self.face_areas[mask] = area(c, k, self.face_positions[mask], self.face_low[mask], self.face_high[mask])
            valid = mask & (self.left >= 0)
            self.offset_left[valid, k] = 0.5 * self.widths[self.left[valid], k]
            valid = mask & (self.right >= 0)
            self.offset_right[valid, k] = -0.5 * self.widths[self.right[valid], k]
        self.interior = (self.left >= 0) & (self.right >= 0)
        if np.any(abs(self.levels[self.left[self.interior]] - self.levels[self.right[self.interior]]) > 1):
            raise ValueError("Adaptive mesh must be face-balanced 2:1")
        self.displacement = self.offset_left - self.offset_right
        lower, upper = self.left < 0, self.right < 0
        self.displacement[lower] = -2 * self.offset_right[lower]
        self.displacement[upper] = 2 * self.offset_left[upper]
        self._gradient_matrix()

    def _gradient_matrix(self):
        delta = self.displacement
        self.weights = 1 / np.sum(delta ** 2, axis=1)
        matrix = np.zeros((len(self.keys), self.dimension, self.dimension))
        part = self.weights[:, None, None] * delta[:, :, None] * delta[:, None, :]
        for idx in (self.left, self.right):
            mask = idx >= 0
            np.add.at(matrix, idx[mask], part[mask])
        self.inverse = np.linalg.inv(matrix)

    def incidence(self):
        """Per-cell face lists in the order the reference accumulates them
        (np.add.at over the faces where the cell is on the left, then over
        those where it is on the right): CSR (start, faces, role) with role
        -1 for a left-cell entry and +1 for a right-cell one."""
        n = len(self.keys)
        fl = np.flatnonzero(self.left >= 0)
        fr = np.flatnonzero(self.right >= 0)
        cells = np.concatenate((self.left[fl], self.right[fr]))
        faces = np.concatenate((fl, fr))
        roles = np.concatenate((np.full(len(fl), -1), np.full(len(fr), 1)))
        order = np.argsort(cells, kind="stable")
        start = np.zeros(n + 1, dtype=np.int64)
        np.add.at(start, cells + 1, 1)
        return np.cumsum(start), faces[order], roles[order]

    def gradients(self, values, left, right, bounded=True):
        """Bounded least-squares slopes (torchAdaptiveGeometry.py:334-372)."""
        d = self.dimension
        diff = right - left
        vector = np.zeros((len(self.keys), d, values.shape[0]))
        part = self.weights[:, None, None] * self.displacement[:, :, None] * diff.T[:, None, :]
        low, high = values.T.copy(), values.T.copy()
        for idx, neighbour in ((self.left, right), (self.right, left)):
            mask = idx >= 0
            np.add.at(vector, idx[mask], part[mask])
            np.minimum.at(low, idx[mask], neighbour.T[mask])
            np.maximum.at(high, idx[mask], neighbour.T[mask])
        grad = np.einsum("nij,njk->nik", self.inverse, vector)
        if bounded:
            theta = np.ones_like(values.T)
            for idx, offset in ((self.left, self.offset_left), (self.right, self.offset_right)):
                mask = idx >= 0
                ids = idx[mask]
                change = np.einsum("nd,ndv->nv", offset[mask], grad[ids])
                allowed = np.where(change >= 0, high[ids] - values.T[ids], low[ids] - values.T[ids])
                ratio = np.ones_like(change)
                np.divide(allowed, change, out=ratio, where=abs(change) > 1e-14)
                np.minimum.at(theta, ids, np.clip(ratio, 0, 1))
            grad *= theta[:, None, :]
        return grad
