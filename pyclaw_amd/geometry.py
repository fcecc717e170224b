"""Grid geometry: Dimension, Grid, Patch, Domain.

Mirrors the reference's geometry API (src/pyclaw/geometry.py:23-666): the same
constructors, attribute names and cell-centre formula
(centers[i] = lower + (i+0.5)*delta, geometry.py:439-446), so problem set-up
code written for the reference runs unchanged.
"""
import math
import os

import numpy as np


def py2_round(v):
    """Python 2's round() to an integer: the nearest integer, halves away from
    zero (the reference is Python 2 code; Python 3's round takes halves to
    even).  v - floor(v) is exact for every finite double, so the test against
    0.5 is exact."""
    if v < 0:
        return -py2_round(-v)
    f = math.floor(v)
    return int(f) + (1 if v - f >= 0.5 else 0)


class Dimension(object):
    """One coordinate direction: Dimension('x', lower, upper, n) or
    Dimension(lower, upper, n, name='x')  (geometry.py:455-493)."""

    def __init__(self, *args, **kwargs):
        self.name = 'x'
        self.units = None
        self.on_lower_boundary = None
        self.on_upper_boundary = None
        if isinstance(args[0], str):
            self.name = args[0]
            self.lower = float(args[1])
            self.upper = float(args[2])
            self.num_cells = int(args[3])
        elif isinstance(args[0], (float, int, np.floating, np.integer)):
            self.lower = float(args[0])
            self.upper = float(args[1])
            self.num_cells = int(args[2])
        else:
            raise Exception("Invalid initializer for Dimension.")
        for k, v in kwargs.items():
            setattr(self, k, v)
        self._centers = None
        self._edges = None

    @property
    def delta(self):
        """Cell width (geometry.py:425-427)."""
        return (self.upper - self.lower) / float(self.num_cells)

    @property
    def edges(self):
        if self._edges is None:
            self._edges = self.lower + np.arange(self.num_cells + 1) * self.delta
        return self._edges

    @property
    def centers(self):
        if self._centers is None:
            self._centers = self.lower + (np.arange(self.num_cells) + 0.5) * self.delta
        return self._centers

    def centers_with_ghost(self, nghost):
        pre = np.linspace(self.lower - (nghost - 0.5) * self.delta, self.lower - 0.5 * self.delta, nghost)
        post = np.linspace(self.upper + 0.5 * self.delta, self.upper + (nghost - 0.5) * self.delta, nghost)
        return np.hstack((pre, self.centers, post))

    def __str__(self):
        out = "Dimension %s" % self.name
        if self.units:
            out += " (%s)" % self.units
        out += ":  (num_cells,delta,[lower,upper]) = (%s,%s,[%s,%s])" % (
            self.num_cells, self.delta, self.lower, self.upper)
        return out


class Grid(object):
    """Dimensions plus derived coordinate arrays and gauges (geometry.py:23-374)."""

    _DIM_ATTRS = ('num_cells', 'lower', 'upper', 'delta', 'units', 'centers', 'edges',
                  'on_lower_boundary', 'on_upper_boundary')

    def __init__(self, dimensions):
        if isinstance(dimensions, Dimension):
            dimensions = [dimensions]
        self._dimensions = []
        self.gauges = []
        self.gauge_file_names = []
        self.gauge_files = []
        self.gauge_dir_name = '_gauges'
        self._c_centers = None
        self._c_edges = None
        for d in dimensions:
            self.add_dimension(d)

    def __getattr__(self, key):
        if key in Grid._DIM_ATTRS:
            return self.get_dim_attribute(key)
        raise AttributeError("'Grid' object has no attribute '" + key + "'")

    @property
    def num_dim(self):
        return len(self._dimensions)

    @property
    def dimensions(self):
        return [getattr(self, n) for n in self._dimensions]

    def add_dimension(self, dimension):
        if dimension.name in self._dimensions:
            raise Exception('Unable to add dimension. A dimension of the same name: %s, already exists.'
                            % dimension.name)
        self._dimensions.append(dimension.name)
        setattr(self, dimension.name, dimension)
        self._c_centers = None
        self._c_edges = None

    def get_dim_attribute(self, attr):
        return [getattr(getattr(self, n), attr) for n in self._dimensions]

    @property
    def c_centers(self):
        if self._c_centers is None:
            if self.num_dim == 1:
                self._c_centers = [self.dimensions[0].centers]
            else:
                idx = np.indices(self.num_cells)
                self._c_centers = [c[idx[i, ...]] for i, c in enumerate(self.get_dim_attribute('centers'))]
        return self._c_centers

    @property
    def c_edges(self):
        if self._c_edges is None:
            if self.num_dim == 1:
                self._c_edges = [self.dimensions[0].edges]
            else:
                idx = np.indices([n + 1 for n in self.num_cells])
                self._c_edges = [e[idx[i, ...]] for i, e in enumerate(self.get_dim_attribute('edges'))]
        return self._c_edges

    p_centers = c_centers
    p_edges = c_edges

    def add_gauges(self, gauge_coords):
        """Cell indices of gauges inside the grid (geometry.py:336-355), rounded as
        Python 2 rounds.  As in the reference, a gauge in the last half cell of a
        dimension gets index num_cells; writing its values then raises IndexError
        (solver.py:722-723)."""
        for gauge in gauge_coords:
            if all(self.lower[n] <= gauge[n] < self.upper[n] for n in range(self.num_dim)):
                idx = [py2_round((gauge[n] - self.lower[n]) / self.delta[n]) for n in range(self.num_dim)]
                self.gauge_file_names.append('gauge' + '_'.join(str(c) for c in gauge) + '.txt')
                self.gauges.append(idx)

    def setup_gauge_files(self, outdir):
        path = os.path.join(outdir, self.gauge_dir_name)
        os.makedirs(path, exist_ok=True)
        for name in self.gauge_file_names:
            f = os.path.join(path, name)
            if os.path.isfile(f):
                os.remove(f)
            self.gauge_files.append(open(f, 'a'))

    def __str__(self):
        return '\n'.join(str(d) for d in self.dimensions)


class Patch(object):
    """A logically rectangular patch (geometry.py:507-598)."""

    def __init__(self, dimensions):
        self.level = 1
        self.patch_index = 1
        if isinstance(dimensions, Dimension):
            dimensions = [dimensions]
        self._dimensions = []
        for d in dimensions:
            d.on_lower_boundary = True
            d.on_upper_boundary = True
            self.add_dimension(d)
        self.grid = Grid(dimensions)

    def add_dimension(self, dimension):
        if dimension.name in self._dimensions:
            raise Exception('Unable to add dimension. A dimension of the same name: %s, already exists.'
                            % dimension.name)
        self._dimensions.append(dimension.name)
        setattr(self, dimension.name, dimension)

    def get_dim_attribute(self, attr):
        return [getattr(getattr(self, n), attr) for n in self._dimensions]

    @property
    def num_cells_global(self):
        return self.get_dim_attribute('num_cells')

    @property
    def lower_global(self):
        return self.get_dim_attribute('lower')

    @property
    def upper_global(self):
        return self.get_dim_attribute('upper')

    @property
    def num_dim(self):
        return len(self._dimensions)

    @property
    def dimensions(self):
        return [getattr(self, n) for n in self._dimensions]

    @property
    def delta(self):
        return self.get_dim_attribute('delta')

    @property
    def name(self):
        return self._dimensions

    def __str__(self):
        return "Patch %s:\n" % self.patch_index + '\n'.join(str(d) for d in self.dimensions)


class Domain(object):
    """A list of patches; Domain(lower, upper, n) or Domain(dimensions)
    (geometry.py:603-666)."""

    def __init__(self, *arg):
        if len(arg) > 1:
            lower, upper, n = arg[0], arg[1], arg[2]
            names = ['x', 'y', 'z'][:len(n)]
            dims = [Dimension(lo, up, nn, name=nm) for lo, up, nn, nm in zip(lower, upper, n, names)]
            self.patches = [Patch(dims)]
        else:
            geom = arg[0]
            if not isinstance(geom, (list, tuple)):
                geom = [geom]
            if isinstance(geom[0], Patch):
                self.patches = list(geom)
            elif isinstance(geom[0], Dimension):
                self.patches = [Patch(list(geom))]
            else:
                raise Exception("Invalid argument list")

    @property
    def num_dim(self):
        return self.patches[0].num_dim

    @property
    def patch(self):
        return self.patches[0]

    @property
    def grid(self):
        return self.patches[0].grid
