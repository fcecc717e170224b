"""Solution: container of States on a Domain (src/pyclaw/solution.py:16-347)."""
import copy

from .geometry import Dimension, Domain, Patch
from .state import State


class Solution(object):
    def __init__(self, *arg, **kwargs):
        self.__dict__['states'] = []
        self.__dict__['domain'] = None
        self.__dict__['_start_frame'] = 0
        if len(arg) == 1:
            frame = arg[0]
            self.__dict__['_start_frame'] = 0 if kwargs.pop('count_from_zero', False) else frame
            self.read(frame, **kwargs)
        elif len(arg) == 2:
            geom = arg[1]
            if isinstance(geom, Domain):
                self.domain = geom
            else:
                geom = list(geom) if isinstance(geom, (list, tuple)) else [geom]
                if isinstance(geom[0], Dimension):
                    self.domain = Domain(Patch(geom))
                elif isinstance(geom[0], Patch):
                    self.domain = Domain(geom)
                else:
                    raise Exception("Invalid argument list")
            s = arg[0]
            if isinstance(s, State):
                self.states.append(s)
            elif isinstance(s, (list, tuple)) and s and isinstance(s[0], State):
                self.states.extend(s)
            elif isinstance(s, int):
                self.states.append(State(self.domain, s))
            if not self.states or self.domain is None:
                raise Exception("Invalid argument list")
        elif len(arg) != 0:
            raise Exception("Invalid argument list")

    def __getattr__(self, key):
        if key in ('t', 'num_eqn', 'mp', 'mF', 'q', 'p', 'F', 'aux', 'capa', 'problem_data', 'num_aux'):
            return getattr(self.states[0], key)
        raise AttributeError("'Solution' object has no attribute '" + key + "'")

    def __setattr__(self, key, value):
        if key in ('t', 'mp', 'mF'):
            for s in self.states:
                setattr(s, key, value)
        else:
            self.__dict__[key] = value

    @property
    def state(self):
        return self.states[0]

    @property
    def patch(self):
        return self.domain.patch

    @property
    def start_frame(self):
        return self._start_frame

    def is_valid(self):
        return all(s.is_valid() for s in self.states)

    def __deepcopy__(self, memo):
        result = Solution()
        result.__dict__['states'] = [copy.deepcopy(s, memo) for s in self.states]
        result.__dict__['domain'] = self.domain
        result.__dict__['_start_frame'] = self._start_frame
        return result

    def write(self, frame, path='./', file_format='ascii', file_prefix=None, write_aux=False,
              options={}, write_p=False):
        from . import io
        io.write(self, frame, path, file_format, file_prefix, write_aux, options, write_p)

    def read(self, frame, path='./_output', file_format='ascii', file_prefix=None, read_aux=True,
             options={}, **kwargs):
        from . import io
        io.read(self, frame, path, file_format, file_prefix, read_aux, options)

    def __str__(self):
        return "states:\n" + ''.join(str(s) for s in self.states)
