from dataclasses import dataclass


@dataclass(frozen=True)
class Bond:
    """A bond connecting two sites (acetn/ipeps/bond.py:4-22).

    k: bond direction (0=left, 1=up, 2=right, 3=down)."""
    s1: tuple
    s2: tuple
    k: int

    def __post_init__(self):
        object.__setattr__(self, 's1', tuple(self.s1))
        object.__setattr__(self, 's2', tuple(self.s2))

    def __iter__(self):
        return iter((self.s1, self.s2, self.k))
