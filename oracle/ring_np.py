"""NumPy restatement of the film ring's slot format — TEST INFRASTRUCTURE ONLY.

* `slot_bytes` — film_producer.py:76-79: x|y f32 of n bodies, a u32 count,
  a fixed-length u32 index field of sub_max entries and sub_max body-major
  f32 tracks (m_sub, 2).
* `schreibe_slot` — film_producer.py:82-121: the position rows of one
  sample (the first 2n of its x|y|vx|vy row), then the complete tracks
  sorted by original index with a stable sort, the lowest sub_max kept,
  the count, the indices and the tracks transposed from (m_sub, 2, nh) to
  (nh, m_sub, 2); returns True on overflow.  Bytes past the count are left
  as they were.
Pinned in tests/test_ring.py against hand-packed slots (struct.pack of every
field) including an overflow whose kept indices are the lowest.
"""
import struct

import numpy as np


def slot_bytes(n: int, m_sub: int, sub_max: int) -> int:
    return 8 * n + 4 + sub_max * (4 + 8 * m_sub)


def schreibe_slot(buf, basis: int, n: int, sub_max: int, out_i, sub_i) -> bool:
    pos_bytes = np.asarray(out_i[:2 * n], "<f4").tobytes()
    buf[basis:basis + 8 * n] = pos_bytes
    off = basis + 8 * n
    if sub_i is None or not sub_max:
        buf[off:off + 4] = struct.pack("<I", 0)
        return False
    idx, pos = sub_i
    keep = np.argsort(idx, kind="stable")[:sub_max]
    nh = len(keep)
    buf[off:off + 4] = struct.pack("<I", nh)
    buf[off + 4:off + 4 + 4 * nh] = np.asarray(idx)[keep].astype("<u4").tobytes()
    off += 4 + 4 * sub_max
    if nh:
        tracks = np.ascontiguousarray(np.asarray(pos)[:, :, keep].transpose(2, 0, 1), "<f4")
        buf[off:off + tracks.nbytes] = tracks.tobytes()
    return len(idx) > sub_max
