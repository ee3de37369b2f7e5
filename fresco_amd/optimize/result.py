"""Result container of the optimizers (interface of fresco/optimize/result.py:10-87:
a dict with attribute access and the legacy tuple protocol
(coords, energy, rms, nfev, self))."""

__all__ = ["Result"]


class Result(dict):
    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError:
            raise AttributeError(name) from None

    __setattr__ = dict.__setitem__
    __delattr__ = dict.__delitem__

    def __repr__(self):
        if self:
            m = max(len(k) for k in self.keys()) + 1
            return "\n".join(k.rjust(m) + ": " + repr(v) for k, v in self.items())
        return self.__class__.__name__ + "()"

    _LEGACY = {0: "coords", 1: "energy", 2: "rms", 3: "nfev"}

    def __getitem__(self, i):
        if isinstance(i, slice):
            return tuple(self[j] for j in range(*i.indices(5)))
        if isinstance(i, int) and not isinstance(i, bool):
            if i == 4:
                return self
            if i in self._LEGACY:
                i = self._LEGACY[i]
        return dict.__getitem__(self, i)

    def __iter__(self):
        return iter((self.coords, self.energy, self.rms, self.nfev, self))
