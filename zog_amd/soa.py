"""Python mirror of ``Zog.SoA`` (lib/zog/soa.ex): the builder that assigns
u32 ids to arbitrary labels in first-seen order and records edges as flat
arrays ready for the native resource.

Behaviour kept from the reference:
  * ids are assigned in first-seen order (ensure_node, soa.ex:85-103);
  * an undirected edge is stored as the arc (u, v) followed by (v, u)
    (add_edge, soa.ex:110-128), so the arc arrays hold both directions;
  * ``to_edge_arrays`` returns the arcs in insertion order (soa.ex:268-279);
  * ``all_labels`` returns labels in insertion order, or ``0..next_id-1``
    for an integer-labelled builder (soa.ex:252-257).
"""
from __future__ import annotations

from typing import Any, Dict, Hashable, Iterable, List, Tuple

import numpy as np


class SoA:
    def __init__(self, kind: str = "directed"):
        if kind not in ("directed", "undirected"):
            raise ValueError("kind must be 'directed' or 'undirected'")
        self.kind = kind
        self.label_to_id: Dict[Hashable, int] = {}
        self.id_to_label: Dict[int, Hashable] = {}
        self._labels: List[Hashable] = []
        self._from: List[int] = []
        self._to: List[int] = []
        self._w: List[float] = []
        self.next_id = 0
        self.integer_labels = False
        # bulk arrays (from_arrays); take precedence over the python lists
        self._arrays = None

    # constructors (soa.ex:48-66)
    @classmethod
    def directed(cls) -> "SoA":
        return cls("directed")

    @classmethod
    def undirected(cls) -> "SoA":
        return cls("undirected")

    @classmethod
    def from_list(cls, kind: str, edges: Iterable[Tuple[Hashable, Hashable, float]]) -> "SoA":
        b = cls(kind)
        for f, t, w in edges:
            b.add_edge(f, t, w)
        return b

    @classmethod
    def from_arrays(cls, kind: str, n: int, src, dst, weights=None) -> "SoA":
        """Integer-labelled builder over ids 0..n-1 with arcs already expanded
        (an undirected graph must already hold both arcs of each edge)."""
        b = cls(kind)
        b.integer_labels = True
        b.next_id = int(n)
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        if weights is None:
            weights = np.ones(len(src), dtype=np.float64)
        b._arrays = (src, dst, np.ascontiguousarray(weights, dtype=np.float64))
        return b

    # nodes
    def ensure_node(self, label: Hashable) -> int:
        if self.integer_labels:
            raise ValueError("integer-labelled builder: use ids directly")
        i = self.label_to_id.get(label)
        if i is None:
            i = self.next_id
            self.label_to_id[label] = i
            self.id_to_label[i] = label
            self._labels.append(label)
            self.next_id += 1
        return i

    def add_node(self, label: Hashable) -> "SoA":
        self.ensure_node(label)
        return self

    # edges
    def add_edge(self, frm: Hashable, to: Hashable, weight: Any = 1.0) -> "SoA":
        if self._arrays is not None:
            raise ValueError("array-built builder is read-only")
        u = self.ensure_node(frm)
        v = self.ensure_node(to)
        w = float(weight)
        self._from.append(u)
        self._to.append(v)
        self._w.append(w)
        if self.kind == "undirected":
            self._from.append(v)
            self._to.append(u)
            self._w.append(w)
        return self

    def add_unweighted_edge(self, frm: Hashable, to: Hashable) -> "SoA":
        return self.add_edge(frm, to, 1.0)

    # queries
    def node_count(self) -> int:
        return self.next_id

    def edge_count(self) -> int:
        return len(self._arrays[0]) if self._arrays is not None else len(self._from)

    def all_labels(self) -> List[Hashable]:
        if self.integer_labels:
            return list(range(self.next_id))
        return list(self._labels)

    def label_of(self, i: int) -> Hashable:
        return i if self.integer_labels else self.id_to_label.get(i)

    def id_of(self, label: Hashable):
        return label if self.integer_labels else self.label_to_id.get(label)

    def to_edge_arrays(self):
        if self._arrays is not None:
            return self._arrays
        return (np.asarray(self._from, dtype=np.uint32), np.asarray(self._to, dtype=np.uint32),
                np.asarray(self._w, dtype=np.float64))
