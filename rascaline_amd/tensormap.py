"""Minimal metatensor-compatible containers (Labels, TensorBlock, TensorMap).

featomic returns metatensor ``TensorMap`` objects (``featomic/src/calculator.rs:414-495``).
metatensor is not available to this package, so this module provides the subset of
its Python API that featomic users and the reference's own tests rely on: label
access, block access, gradients, and the ``keys_to_samples`` /
``keys_to_properties`` / ``components_to_properties`` transformations used to
flatten descriptors (``featomic/tests/soap-spherical-expansion.rs:17-24``).
"""

from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np


class Labels:
    """Named integer labels, one row per entry (metatensor ``Labels``)."""

    def __init__(self, names: Sequence[str], values):
        self.names = list(names)
        values = np.asarray(values, dtype=np.int32)
        if values.ndim == 1 and len(self.names) == 1:
            values = values.reshape(-1, 1)
        if values.ndim != 2:
            values = values.reshape(-1, len(self.names)) if len(self.names) else values.reshape(len(values), 0)
        self.values = values

    @staticmethod
    def range(name: str, n: int) -> "Labels":
        return Labels([name], np.arange(n, dtype=np.int32).reshape(-1, 1))

    @staticmethod
    def empty(names) -> "Labels":
        return Labels(names, np.zeros((0, len(names)), dtype=np.int32))

    def __len__(self):
        return self.values.shape[0]

    def __getitem__(self, i):
        return self.values[i]

    def column(self, name: str) -> np.ndarray:
        return self.values[:, self.names.index(name)]

    def position(self, entry) -> Optional[int]:
        entry = np.asarray(entry, dtype=np.int32)
        hits = np.nonzero(np.all(self.values == entry, axis=1))[0]
        return int(hits[0]) if len(hits) else None

    def __eq__(self, other):
        return (
            isinstance(other, Labels)
            and self.names == other.names
            and self.values.shape == other.values.shape
            and bool(np.all(self.values == other.values))
        )

    def __repr__(self):
        return f"Labels({self.names}, {len(self)} entries)"


class TensorBlock:
    def __init__(self, values, samples: Labels, components: List[Labels], properties: Labels):
        self.values = values
        self.samples = samples
        self.components = list(components)
        self.properties = properties
        self._gradients: Dict[str, "TensorBlock"] = {}

    def add_gradient(self, parameter: str, gradient: "TensorBlock"):
        self._gradients[parameter] = gradient

    def gradient(self, parameter: str) -> "TensorBlock":
        return self._gradients[parameter]

    def gradients_list(self) -> List[str]:
        return list(self._gradients.keys())

    def __repr__(self):
        return f"TensorBlock(values={tuple(self.values.shape)}, gradients={self.gradients_list()})"


class TensorMap:
    def __init__(self, keys: Labels, blocks: List[TensorBlock]):
        assert len(keys) == len(blocks)
        self.keys = keys
        self.blocks_ = list(blocks)

    def __len__(self):
        return len(self.blocks_)

    def blocks(self):
        return list(self.blocks_)

    def block_by_id(self, i: int) -> TensorBlock:
        return self.blocks_[i]

    def block(self, selection=None, **kwargs) -> TensorBlock:
        if selection is None and not kwargs:
            if len(self.blocks_) != 1:
                raise ValueError("block() without selection requires a single block")
            return self.blocks_[0]
        if selection is None:
            selection = kwargs
        if isinstance(selection, int):
            return self.blocks_[selection]
        mask = np.ones(len(self.keys), dtype=bool)
        for name, value in dict(selection).items():
            mask &= self.keys.column(name) == value
        hits = np.nonzero(mask)[0]
        if len(hits) != 1:
            raise ValueError(f"selection {selection} matches {len(hits)} blocks")
        return self.blocks_[int(hits[0])]

    def items(self):
        return list(zip(self.keys.values, self.blocks_))

    # ------------------------------------------------------------------ #
    # metatensor-style transformations

    def _group(self, names):
        move_idx = [self.keys.names.index(n) for n in names]
        keep_idx = [i for i in range(len(self.keys.names)) if i not in move_idx]
        groups: Dict[tuple, List[int]] = {}
        order = []
        for b, key in enumerate(self.keys.values):
            kept = tuple(int(key[i]) for i in keep_idx)
            if kept not in groups:
                groups[kept] = []
                order.append(kept)
            groups[kept].append(b)
        keep_names = [self.keys.names[i] for i in keep_idx]
        return move_idx, keep_names, order, groups

    def keys_to_samples(self, names, sort_samples: bool = True) -> "TensorMap":
        if isinstance(names, str):
            names = [names]
        move_idx, keep_names, order, groups = self._group(names)
        new_blocks = []
        for kept in order:
            members = groups[kept]
            first = self.blocks_[members[0]]
            rows = []
            for b in members:
                blk = self.blocks_[b]
                moved = np.array([self.keys.values[b][i] for i in move_idx], dtype=np.int32)
                for s, sv in enumerate(blk.samples.values):
                    rows.append((tuple(int(x) for x in sv) + tuple(int(x) for x in moved), b, s))
            if sort_samples:
                rows.sort(key=lambda r: r[0])
            n_new = len(rows)
            samples = Labels(first.samples.names + list(names), np.array([r[0] for r in rows], dtype=np.int32).reshape(n_new, -1))
            values = np.zeros((n_new,) + first.values.shape[1:], dtype=first.values.dtype)
            remap = {}
            for new_i, (_, b, s) in enumerate(rows):
                values[new_i] = self.blocks_[b].values[s]
                remap[(b, s)] = new_i
            block = TensorBlock(values, samples, first.components, first.properties)
            for param in first.gradients_list():
                grows = []
                for b in members:
                    g = self.blocks_[b].gradient(param)
                    for r, sv in enumerate(g.samples.values):
                        new_s = remap[(b, int(sv[0]))]
                        grows.append(((new_s,) + tuple(int(x) for x in sv[1:]), b, r))
                if sort_samples:
                    grows.sort(key=lambda r: r[0])
                g0 = first.gradient(param)
                gvals = np.zeros((len(grows),) + g0.values.shape[1:], dtype=g0.values.dtype)
                for k, (_, b, r) in enumerate(grows):
                    gvals[k] = self.blocks_[b].gradient(param).values[r]
                gs = Labels(g0.samples.names, np.array([r[0] for r in grows], dtype=np.int32).reshape(len(grows), -1))
                block.add_gradient(param, TensorBlock(gvals, gs, g0.components, g0.properties))
            new_blocks.append(block)
        return TensorMap(Labels(keep_names, np.array(order, dtype=np.int32).reshape(len(order), -1)), new_blocks)

    def keys_to_properties(self, names, sort_samples: bool = True) -> "TensorMap":
        if isinstance(names, str):
            names = [names]
        move_idx, keep_names, order, groups = self._group(names)
        new_blocks = []
        for kept in order:
            members = groups[kept]
            first = self.blocks_[members[0]]
            # union of samples
            sample_set = {}
            sample_list = []
            for b in members:
                for sv in self.blocks_[b].samples.values:
                    t = tuple(int(x) for x in sv)
                    if t not in sample_set:
                        sample_set[t] = None
                        sample_list.append(t)
            if sort_samples:
                sample_list.sort()
            sample_index = {t: i for i, t in enumerate(sample_list)}
            prop_rows = []
            prop_slices = []
            start = 0
            for b in members:
                blk = self.blocks_[b]
                moved = tuple(int(self.keys.values[b][i]) for i in move_idx)
                for pv in blk.properties.values:
                    prop_rows.append(moved + tuple(int(x) for x in pv))
                prop_slices.append((start, start + len(blk.properties)))
                start += len(blk.properties)
            properties = Labels(list(names) + first.properties.names, np.array(prop_rows, dtype=np.int32).reshape(len(prop_rows), -1))
            shape = (len(sample_list),) + first.values.shape[1:-1] + (len(prop_rows),)
            values = np.zeros(shape, dtype=first.values.dtype)
            for b, (p0, p1) in zip(members, prop_slices):
                blk = self.blocks_[b]
                idx = [sample_index[tuple(int(x) for x in sv)] for sv in blk.samples.values]
                if len(idx):
                    values[idx, ..., p0:p1] = blk.values
            samples = Labels(first.samples.names, np.array(sample_list, dtype=np.int32).reshape(len(sample_list), -1))
            block = TensorBlock(values, samples, first.components, properties)
            for param in first.gradients_list():
                grad_set = {}
                grad_list = []
                for b in members:
                    blk = self.blocks_[b]
                    g = blk.gradient(param)
                    for sv in g.samples.values:
                        old_sample = tuple(int(x) for x in blk.samples.values[int(sv[0])])
                        t = (sample_index[old_sample],) + tuple(int(x) for x in sv[1:])
                        if t not in grad_set:
                            grad_set[t] = None
                            grad_list.append(t)
                if sort_samples:
                    grad_list.sort()
                grad_index = {t: i for i, t in enumerate(grad_list)}
                g0 = first.gradient(param)
                gshape = (len(grad_list),) + g0.values.shape[1:-1] + (len(prop_rows),)
                gvals = np.zeros(gshape, dtype=g0.values.dtype)
                for b, (p0, p1) in zip(members, prop_slices):
                    blk = self.blocks_[b]
                    g = blk.gradient(param)
                    for r, sv in enumerate(g.samples.values):
                        old_sample = tuple(int(x) for x in blk.samples.values[int(sv[0])])
                        t = (sample_index[old_sample],) + tuple(int(x) for x in sv[1:])
                        gvals[grad_index[t], ..., p0:p1] = g.values[r]
                gs = Labels(g0.samples.names, np.array(grad_list, dtype=np.int32).reshape(len(grad_list), -1))
                block.add_gradient(param, TensorBlock(gvals, gs, g0.components, properties))
            new_blocks.append(block)
        return TensorMap(Labels(keep_names, np.array(order, dtype=np.int32).reshape(len(order), -1)), new_blocks)

    def components_to_properties(self, names) -> "TensorMap":
        if isinstance(names, str):
            names = [names]
        new_blocks = []
        for blk in self.blocks_:
            ci = [c.names[0] for c in blk.components].index(names[0])
            comp = blk.components[ci]
            rest = [c for k, c in enumerate(blk.components) if k != ci]
            nc, npr = len(comp), len(blk.properties)
            rows = []
            for cv in comp.values:
                for pv in blk.properties.values:
                    rows.append(tuple(int(x) for x in cv) + tuple(int(x) for x in pv))
            properties = Labels(comp.names + blk.properties.names, np.array(rows, dtype=np.int32).reshape(len(rows), -1))
            v = np.moveaxis(blk.values, 1 + ci, -2)
            v = v.reshape(v.shape[:-2] + (nc * npr,))
            block = TensorBlock(np.ascontiguousarray(v), blk.samples, rest, properties)
            for param in blk.gradients_list():
                g = blk.gradient(param)
                n_extra = len(g.components) - len(blk.components)
                gv = np.moveaxis(g.values, 1 + n_extra + ci, -2)
                gv = gv.reshape(gv.shape[:-2] + (nc * npr,))
                gcomps = g.components[:n_extra] + rest
                block.add_gradient(param, TensorBlock(np.ascontiguousarray(gv), g.samples, gcomps, properties))
            new_blocks.append(block)
        return TensorMap(self.keys, new_blocks)

    def __repr__(self):
        return f"TensorMap({len(self)} blocks, keys={self.keys.names})"
