"""First-fit range allocator for the shared case pools (host logic)."""

import bisect


class RangeAllocator:
    """Hands out [offset, offset + size) ranges of [0, capacity).  Free ranges
    are kept sorted and coalesced; allocation is first fit, rounded up to
    `align`."""

    def __init__(self, capacity: int, align: int = 1):
        self.capacity = int(capacity)
        self.align = int(align)
        self._starts = [0]
        self._sizes = [self.capacity]
        self.used = 0
        self.high_water = 0

    def _round(self, n):
        a = self.align
        return ((int(n) + a - 1) // a) * a

    def alloc(self, size: int):
        size = self._round(size)
        if size <= 0:
            raise ValueError("allocation size must be positive")
        for i, (s, n) in enumerate(zip(self._starts, self._sizes)):
            if n >= size:
                if n == size:
                    del self._starts[i]
                    del self._sizes[i]
                else:
                    self._starts[i] = s + size
                    self._sizes[i] = n - size
                self.used += size
                self.high_water = max(self.high_water, s + size)
                return s
        return None

    def free(self, offset: int, size: int):
        size = self._round(size)
        i = bisect.bisect_left(self._starts, offset)
        if i < len(self._starts) and self._starts[i] < offset + size:
            raise ValueError("double free or overlapping free")
        if i > 0 and self._starts[i - 1] + self._sizes[i - 1] > offset:
            raise ValueError("double free or overlapping free")
        self._starts.insert(i, offset)
        self._sizes.insert(i, size)
        self.used -= size
        # coalesce with the right and the left neighbour
        if i + 1 < len(self._starts) and self._starts[i] + self._sizes[i] == self._starts[i + 1]:
            self._sizes[i] += self._sizes[i + 1]
            del self._starts[i + 1]
            del self._sizes[i + 1]
        if i > 0 and self._starts[i - 1] + self._sizes[i - 1] == self._starts[i]:
            self._sizes[i - 1] += self._sizes[i]
            del self._starts[i]
            del self._sizes[i]
        # the high-water mark drops when the top range is free again
        if self._starts and self._starts[-1] + self._sizes[-1] == self.capacity:
            self.high_water = min(self.high_water, self._starts[-1])

    def fits(self, size: int) -> bool:
        size = self._round(size)
        return any(n >= size for n in self._sizes)
