"""SquareLattice restated (TeNeS src/SquareLattice.hpp:54-163,
src/SquareLattice.cpp:48-137): site index x + y*LX, neighbour directions
0:left 1:top 2:right 3:bottom, skew boundary T(x, y) = T(x + skew, y + LY)."""
from __future__ import annotations

from math import gcd


class SquareLattice:
    def __init__(self, LX: int, LY: int, skew: int = 0):
        if LX <= 0 or LY <= 0:
            raise ValueError("Lattice.X and Lattice.Y should be positive")
        self.LX, self.LY = LX, LY
        self.N_UNIT = LX * LY
        self.skew = skew % LX
        self.LX_noskew = LX
        if self.skew == 0:
            self.LY_noskew = LY
        else:
            lcm = LX // gcd(LX, self.skew) * self.skew
            self.LY_noskew = LY * (lcm // self.skew)
        self.N_UNIT_noskew = self.LX_noskew * self.LY_noskew
        N = self.N_UNIT
        nn = [[0] * 4 for _ in range(N)]
        for iy in range(LY):
            for ix in range(LX):
                i = ix + iy * LX
                nn[i][0] = (ix - 1 + LX) % LX + iy * LX
                nn[i][1] = ix + ((iy + 1) % LY) * LX
                nn[i][2] = (ix + 1) % LX + iy * LX
                nn[i][3] = ix + ((iy - 1 + LY) % LY) * LX
        if self.skew != 0:
            for ix in range(LX):
                nn[ix][3] = (ix + LX + self.skew) % LX + (LY - 1) * LX
                nn[N + ix - LX][1] = (ix + LX - self.skew) % LX
        self._nn = nn
        par = [0] * N
        for ix in range(LX):
            par[ix] = -1 if ix % 2 != 0 else 1
        for iy in range(1, LY):
            for ix in range(LX):
                s = self.index(ix, iy)
                par[s] = -par[self.bottom(s)]
        self._par = par

    def x(self, i: int) -> int:
        return i % self.LX

    def y(self, i: int) -> int:
        return i // self.LX

    def index(self, x: int, y: int) -> int:
        y_offset = 0
        if y >= self.LY:
            y_offset = y // self.LY
        elif y < 0:
            # C++ integer division truncates toward zero
            y_offset = int((y + 1) / self.LY) - 1
        y -= self.LY * y_offset
        x -= self.skew * y_offset
        x %= self.LX
        return x + y * self.LX

    def neighbor(self, i: int, d: int) -> int:
        return self._nn[i][d]

    def left(self, i):
        return self._nn[i][0]

    def top(self, i):
        return self._nn[i][1]

    def right(self, i):
        return self._nn[i][2]

    def bottom(self, i):
        return self._nn[i][3]

    def parity(self, i):
        return self._par[i]

    def other(self, i: int, dx: int, dy: int) -> int:
        while dx != 0:
            if dx > 0:
                i = self.right(i)
                dx -= 1
            else:
                i = self.left(i)
                dx += 1
        while dy != 0:
            if dy > 0:
                i = self.top(i)
                dy -= 1
            else:
                i = self.bottom(i)
                dy += 1
        return i
