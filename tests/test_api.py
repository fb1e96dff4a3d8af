"""The user API's validation and properties, as the reference's
test/unit/test_api.py states them (:74-840): argument types and shapes are
checked at construction with the reference's exception classes, data are
converted and reshaped the same way, equality of Sets is identity and of
Maps is attribute equality, and illegal iteration-set/map combinations are
caught when a ParLoop is built.  Host logic only: runs on the oracle and,
marked gpu, on hip.  (numpy 2 has no ``np.int``: Python ints become the
platform default integer, ``np.dtype(int)``.)"""

import subprocess
import sys

import numpy as np
import pytest

from pyop2_amd import op2, base, exceptions
from pyop2_amd import backends as _backends
from pyop2_amd import configuration as cfg


@pytest.fixture(params=[1, 2, (2, 3)])
def set(request):
    return op2.Set(5, request.param, 'foo')


@pytest.fixture
def iterset():
    return op2.Set(2, 1, 'iterset')


@pytest.fixture
def dataset():
    return op2.Set(3, 1, 'dataset')


@pytest.fixture
def m(iterset, dataset):
    return op2.Map(iterset, dataset, 2, [1] * 2 * iterset.size, 'm')


@pytest.fixture
def const(request):
    c = op2.Const(1, 1, 'test_const_nonunique_name')
    request.addfinalizer(c.remove_from_namespace)
    return c


@pytest.fixture
def sparsity(m):
    return op2.Sparsity((m, m))


def _size(s):
    return s.size * int(np.prod(s.dim))


class TestInitAPI:

    def test_noninit(self):
        """Constructing op2 objects before init raises RuntimeError (in a
        fresh interpreter: this one has a backend already)."""
        code = ("import sys; sys.path.insert(0, %r)\n"
                "from pyop2_amd import op2\n"
                "try:\n    op2.Set(1)\nexcept RuntimeError:\n    sys.exit(0)\n"
                "sys.exit(1)\n" % sys.path[0])
        assert subprocess.run([sys.executable, '-c', code]).returncode == 0

    def test_invalid_init(self):
        code = ("import sys; sys.path.insert(0, %r)\n"
                "from pyop2_amd import op2\n"
                "try:\n    op2.init(backend='invalid_backend')\nexcept ImportError:\n"
                "    sys.exit(0)\nsys.exit(1)\n" % sys.path[0])
        assert subprocess.run([sys.executable, '-c', code]).returncode == 0

    def test_init(self, backend):
        assert _backends.get_backend() == backend

    def test_double_init(self, backend):
        """init again with the same backend updates the configuration."""
        op2.init(backend=backend, foo='bar')
        assert _backends.get_backend() == backend
        assert cfg.get('foo') == 'bar'

    def test_change_backend_fails(self, backend):
        with pytest.raises(RuntimeError):
            op2.init(backend='other')


class TestMPIAPI:

    def test_running_sequentially(self, backend):
        assert not op2.MPI.parallel
        assert op2.MPI.comm.size == 1 and op2.MPI.comm.rank == 0


class TestAccessAPI:

    @pytest.mark.parametrize("mode", base.Access._modes)
    def test_access(self, backend, mode):
        assert repr(base.Access(mode)) == "Access('%s')" % mode

    def test_illegal_access(self, backend):
        with pytest.raises(exceptions.ModeValueError):
            base.Access('ILLEGAL_ACCESS')


class TestSetAPI:

    def test_set_illegal_size(self, backend):
        with pytest.raises(exceptions.SizeTypeError):
            op2.Set('illegalsize')

    def test_set_illegal_dim(self, backend):
        with pytest.raises(TypeError):
            op2.Set(1, 'illegaldim')

    def test_set_illegal_dim_tuple(self, backend):
        with pytest.raises(TypeError):
            op2.Set(1, (1, 'illegaldim'))

    def test_set_illegal_name(self, backend):
        with pytest.raises(exceptions.NameTypeError):
            op2.Set(1, 1, 2)

    def test_set_dim(self, backend):
        assert op2.Set(1, 1).dim == (1,)

    def test_set_dim_list(self, backend):
        assert op2.Set(1, [2, 3]).dim == (2, 3)

    def test_set_repr_str(self, backend, set):
        assert repr(set) and str(set)

    def test_set_equality(self, backend, set):
        """Set equality is identity."""
        copy = op2.Set(set.size, set.dim, set.name)
        assert set == set and set != copy


class TestDatAPI:

    def test_dat_illegal_set(self, backend):
        with pytest.raises(exceptions.SetTypeError):
            op2.Dat('illegalset', 1)

    def test_dat_illegal_name(self, backend, set):
        with pytest.raises(exceptions.NameTypeError):
            op2.Dat(set, name=2)

    def test_dat_initialise_data(self, backend, set):
        d = op2.Dat(set)
        assert d.data.size == _size(set) and d.data.dtype == np.float64

    def test_dat_initialise_data_type(self, backend, set):
        assert op2.Dat(set, dtype=np.int32).data.dtype == np.int32

    def test_dat_illegal_map(self, backend, set):
        """A Dat called with a map whose target is another Set."""
        d = op2.Dat(set)
        to_set2 = op2.Map(op2.Set(3), op2.Set(2), 1, [0, 0, 0])
        with pytest.raises(exceptions.MapValueError):
            d(to_set2, op2.READ)

    def test_dat_dtypes(self, backend, set):
        n = _size(set)
        assert op2.Dat(set).dtype == np.float64
        assert op2.Dat(set, [1.0] * n).dtype == np.float64
        assert op2.Dat(set, [1] * n).dtype == np.dtype(int)
        assert op2.Dat(set, [1] * n, np.double).dtype == np.float64
        assert op2.Dat(set, [1.5] * n, np.int32).dtype == np.int32

    def test_dat_illegal_dtype(self, backend, set):
        with pytest.raises(exceptions.DataTypeError):
            op2.Dat(set, dtype='illegal_type')

    def test_dat_illegal_length(self, backend, set):
        with pytest.raises(exceptions.DataValueError):
            op2.Dat(set, [1] * (_size(set) + 1))

    def test_dat_reshape(self, backend, set):
        d = op2.Dat(set, [1.0] * _size(set))
        assert d.data.shape == (set.size,) + set.dim

    def test_dat_properties(self, backend, set):
        d = op2.Dat(set, [1] * _size(set), 'double', 'bar')
        assert d.dataset == set and d.dtype == np.float64 and d.name == 'bar' \
            and d.data.sum() == _size(set)

    def test_dat_ro_accessor(self, backend, set):
        d = op2.Dat(set, range(_size(set)), dtype=np.int32)
        x = d.data_ro
        with pytest.raises((RuntimeError, ValueError)):
            x[0] = 1

    def test_dat_ro_write_accessor(self, backend, set):
        d = op2.Dat(set, range(_size(set)), dtype=np.int32)
        x = d.data_ro
        with pytest.raises((RuntimeError, ValueError)):
            x[0] = 1
        x = d.data
        x[0] = -100
        assert (d.data_ro[0] == -100).all()


class TestSparsityAPI:

    @pytest.fixture
    def mi(self, dataset):
        it = op2.Set(3, 1, 'iterset2')
        return op2.Map(it, dataset, 1, [1] * it.size, 'mi')

    @pytest.fixture
    def md(self, iterset):
        ds = op2.Set(1, 1, 'dataset2')
        return op2.Map(iterset, ds, 1, [0] * iterset.size, 'md')

    def test_sparsity_illegal_rmap(self, backend, m):
        with pytest.raises(TypeError):
            op2.Sparsity(('illegalrmap', m))

    def test_sparsity_illegal_cmap(self, backend, m):
        with pytest.raises(TypeError):
            op2.Sparsity((m, 'illegalcmap'))

    def test_sparsity_single_map(self, backend, m):
        s = op2.Sparsity(m, "foo")
        assert s.maps[0] == (m, m) and s.dims == (1, 1) and s.name == "foo"

    def test_sparsity_map_pair(self, backend, m):
        s = op2.Sparsity((m, m), "foo")
        assert s.maps[0] == (m, m) and s.dims == (1, 1) and s.name == "foo"

    def test_sparsity_map_pair_different_dataset(self, backend, m, md):
        s = op2.Sparsity((m, md), "foo")
        assert s.maps[0] == (m, md) and s.dims == (1, 1) and s.name == "foo"

    def test_sparsity_multiple_map_pairs(self, backend, m):
        s = op2.Sparsity(((m, m), (m, m)), "foo")
        assert list(s.maps) == [(m, m), (m, m)] and s.dims == (1, 1)

    def test_sparsity_map_pairs_different_itset(self, backend, m, mi):
        s = op2.Sparsity(((m, m), (mi, mi)), "foo")
        assert len(s.maps) == 2 and s.dims == (1, 1)

    def test_sparsity_illegal_itersets(self, backend, m, mi):
        with pytest.raises(RuntimeError):
            op2.Sparsity((m, mi))

    def test_sparsity_illegal_row_datasets(self, backend, m, md):
        with pytest.raises(RuntimeError):
            op2.Sparsity(((m, m), (md, m)))

    def test_sparsity_illegal_col_datasets(self, backend, m, md):
        with pytest.raises(RuntimeError):
            op2.Sparsity(((m, m), (m, md)))


class TestMatAPI:

    def test_mat_illegal_sets(self, backend):
        with pytest.raises(TypeError):
            op2.Mat('illegalsparsity')

    def test_mat_illegal_name(self, backend, sparsity):
        with pytest.raises(exceptions.NameTypeError):
            op2.Mat(sparsity, name=2)

    def test_mat_dtype(self, backend, sparsity):
        assert op2.Mat(sparsity).dtype == np.float64

    def test_mat_properties(self, backend, sparsity):
        mat = op2.Mat(sparsity, 'double', 'bar')
        assert mat.sparsity == sparsity and mat.dtype == np.float64 and mat.name == 'bar'

    def test_mat_illegal_maps(self, backend, sparsity):
        mat = op2.Mat(sparsity)
        wrongmap = op2.Map(op2.Set(2), op2.Set(3), 2, [0, 0, 0, 0])
        with pytest.raises(exceptions.MapValueError):
            mat((wrongmap[0], wrongmap[1]), op2.INC)


class TestConstAPI:

    def test_const_illegal_dim(self, backend):
        with pytest.raises(TypeError):
            op2.Const('illegaldim', 1, 'test_const_illegal_dim')

    def test_const_illegal_dim_tuple(self, backend):
        with pytest.raises(TypeError):
            op2.Const((1, 'illegaldim'), 1, 'test_const_illegal_dim_tuple')

    def test_const_nonunique_name(self, backend, const):
        with pytest.raises(op2.Const.NonUniqueNameError):
            op2.Const(1, 1, 'test_const_nonunique_name')

    def test_const_remove_from_namespace(self, backend):
        c = op2.Const(1, 1, 'test_const_remove_from_namespace')
        c.remove_from_namespace()
        c = op2.Const(1, 1, 'test_const_remove_from_namespace')
        c.remove_from_namespace()
        assert c.name == 'test_const_remove_from_namespace'

    def test_const_illegal_name(self, backend):
        with pytest.raises(exceptions.NameTypeError):
            op2.Const(1, 1, 2)

    def test_const_dims_and_dtypes(self, backend):
        cases = [((1, 1, 'c_dim'), (1,), None),
                 (([2, 3], [1] * 6, 'c_dim_list'), (2, 3), None),
                 ((1, 1.0, 'c_float'), None, np.float64),
                 ((1, 1, 'c_int'), None, np.dtype(int)),
                 ((1, 1, 'c_int_float', 'double'), None, np.float64),
                 ((1, 1.5, 'c_float_int', 'int'), None, np.dtype(int))]
        for args, dim, dtype in cases:
            c = op2.Const(*args)
            c.remove_from_namespace()
            if dim is not None:
                assert c.dim == dim
            if dtype is not None:
                assert c.dtype == dtype

    def test_const_illegal_dtype(self, backend):
        with pytest.raises(exceptions.DataValueError):
            op2.Const(1, 'illegal_type', 'test_const_illegal_dtype', 'double')

    @pytest.mark.parametrize("dim", [1, (2, 2)])
    def test_const_illegal_length(self, backend, dim):
        with pytest.raises(exceptions.DataValueError):
            op2.Const(dim, [1] * (int(np.prod(dim)) + 1),
                      'test_const_illegal_length_%r' % int(np.prod(dim)))

    def test_const_reshape_and_properties(self, backend):
        c = op2.Const((2, 2), [1] * 4, 'baz', 'double')
        c.remove_from_namespace()
        assert c.dim == (2, 2) and c.data.shape == (2, 2) and c.dtype == np.float64 \
            and c.name == 'baz' and c.data.sum() == 4

    def test_const_setter(self, backend):
        c = op2.Const(1, 1, 'c')
        c.remove_from_namespace()
        c.data = 2
        assert c.data.sum() == 2
        with pytest.raises(exceptions.DataValueError):
            c.data = [1, 2]


class TestGlobalAPI:

    def test_global_illegal_dim(self, backend):
        with pytest.raises(TypeError):
            op2.Global('illegaldim')

    def test_global_illegal_dim_tuple(self, backend):
        with pytest.raises(TypeError):
            op2.Global((1, 'illegaldim'))

    def test_global_illegal_name(self, backend):
        with pytest.raises(exceptions.NameTypeError):
            op2.Global(1, 1, name=2)

    def test_global_dims_and_dtypes(self, backend):
        assert op2.Global(1, 1).dim == (1,)
        assert op2.Global([2, 3], [1] * 6).dim == (2, 3)
        assert op2.Global(1, 1.0).dtype == np.float64
        assert op2.Global(1, 1).dtype == np.dtype(int)
        assert op2.Global(1, 1, 'double').dtype == np.float64
        assert op2.Global(1, 1.5, 'int').dtype == np.dtype(int)

    def test_global_illegal_dtype(self, backend):
        with pytest.raises(exceptions.DataValueError):
            op2.Global(1, 'illegal_type', 'double')

    @pytest.mark.parametrize("dim", [1, (2, 2)])
    def test_global_illegal_length(self, backend, dim):
        with pytest.raises(exceptions.DataValueError):
            op2.Global(dim, [1] * (int(np.prod(dim)) + 1))

    def test_global_reshape_and_properties(self, backend):
        g = op2.Global((2, 2), [1] * 4, 'double', 'bar')
        assert g.dim == (2, 2) and g.data.shape == (2, 2) and g.dtype == np.float64 \
            and g.name == 'bar' and g.data.sum() == 4

    def test_global_setter(self, backend):
        g = op2.Global(1, 1)
        g.data = 2
        assert g.data.sum() == 2
        with pytest.raises(exceptions.DataValueError):
            g.data = [1, 2]


class TestMapAPI:

    def test_map_illegal_iterset(self, backend, set):
        with pytest.raises(exceptions.SetTypeError):
            op2.Map('illegalset', set, 1, [])

    def test_map_illegal_dataset(self, backend, set):
        with pytest.raises(exceptions.SetTypeError):
            op2.Map(set, 'illegalset', 1, [])

    def test_map_illegal_dim(self, backend, set):
        with pytest.raises(exceptions.DimTypeError):
            op2.Map(set, set, 'illegaldim', [])

    def test_map_illegal_dim_tuple(self, backend, set):
        with pytest.raises(exceptions.DimTypeError):
            op2.Map(set, set, (2, 2), [])

    def test_map_illegal_name(self, backend, set):
        with pytest.raises(exceptions.NameTypeError):
            op2.Map(set, set, 1, [], name=2)

    def test_map_illegal_dtype(self, backend, set):
        with pytest.raises(exceptions.DataValueError):
            op2.Map(set, set, 1, 'abcdefg')

    def test_map_illegal_length(self, backend, iterset, dataset):
        with pytest.raises(exceptions.DataValueError):
            op2.Map(iterset, dataset, 1, [1] * (iterset.size + 1))

    def test_map_convert_float_int(self, backend, iterset, dataset):
        mp = op2.Map(iterset, dataset, 1, [1.5] * iterset.size)
        assert mp.values.dtype == np.int32 and mp.values.sum() == iterset.size

    def test_map_reshape_and_properties(self, backend, iterset, dataset):
        mp = op2.Map(iterset, dataset, 2, [1] * 2 * iterset.size, 'bar')
        assert mp.values.shape == (iterset.size, 2)
        assert mp.iterset == iterset and mp.dataset == dataset and mp.dim == 2 \
            and mp.values.sum() == 2 * iterset.size and mp.name == 'bar'

    def test_map_indexing(self, backend, m):
        assert m[0].idx == 0

    def test_map_slicing(self, backend, m):
        with pytest.raises(NotImplementedError):
            m[:]

    def test_map_equality(self, backend, m):
        """Maps are equal when their attributes are (Sets by identity)."""
        assert m == op2.Map(m.iterset, m.dataset, m.dim, m.values, m.name)
        itercopy = op2.Set(m.iterset.size, m.iterset.dim, m.iterset.name)
        assert m != op2.Map(itercopy, m.dataset, m.dim, m.values, m.name)
        assert m != op2.Map(m.iterset, m.dataset, m.dim * 2, list(m.values) * 2, m.name)
        assert m != op2.Map(m.iterset, m.dataset, m.dim, m.values, 'n')


class TestIterationSpaceAPI:

    def test_iteration_space_illegal_iterset(self, backend):
        with pytest.raises(exceptions.SetTypeError):
            op2.IterationSpace('illegalset', 1)

    def test_iteration_space_illegal_extents(self, backend, set):
        with pytest.raises(TypeError):
            op2.IterationSpace(set, 'illegalextents')

    def test_iteration_space_illegal_extents_tuple(self, backend, set):
        with pytest.raises(TypeError):
            op2.IterationSpace(set, (1, 'illegalextents'))

    def test_iteration_space_extents(self, backend, set):
        assert op2.IterationSpace(set, 1).extents == (1,)
        assert op2.IterationSpace(set, [2, 3]).extents == (2, 3)
        i = op2.IterationSpace(set, (2, 3))
        assert i.iterset == set and i.extents == (2, 3)


class TestKernelAPI:

    def test_kernel_illegal_name(self, backend):
        with pytest.raises(exceptions.NameTypeError):
            op2.Kernel("", name=2)

    def test_kernel_properties(self, backend):
        assert op2.Kernel("", 'foo').name == 'foo'


class TestIllegalItersetMaps:

    def test_illegal_dat_iterset(self, backend):
        set1, set2 = op2.Set(2), op2.Set(3)
        dat = op2.Dat(set1)
        mp = op2.Map(set2, set1, 1, [0, 0, 0])
        kernel = op2.Kernel("void k() { }", "k")
        with pytest.raises(exceptions.MapValueError):
            base.ParLoop(kernel, set1, dat(mp, op2.READ))

    def test_illegal_mat_iterset(self, backend, sparsity):
        set1 = op2.Set(2)
        mat = op2.Mat(sparsity)
        rmap, cmap = sparsity.maps[0]
        kernel = op2.Kernel("void k() { }", "k")
        with pytest.raises(exceptions.MapValueError):
            base.ParLoop(kernel, set1(3, 3), mat((rmap[op2.i[0]], cmap[op2.i[1]]), op2.INC))


class TestSolverAPI:

    def test_solver_defaults(self, backend):
        assert op2.Solver().parameters == base.DEFAULT_SOLVER_PARAMETERS

    def test_set_options(self, backend):
        params = {'linear_solver': 'gmres', 'maximum_iterations': 25}
        for s in (op2.Solver(params), op2.Solver(linear_solver='gmres', maximum_iterations=25)):
            assert s.parameters['linear_solver'] == 'gmres' \
                and s.parameters['maximum_iterations'] == 25
        s = op2.Solver()
        s.update_parameters(params)
        assert s.parameters['linear_solver'] == 'gmres' \
            and s.parameters['maximum_iterations'] == 25

    def test_set_params_and_kwargs_illegal(self, backend):
        with pytest.raises(RuntimeError):
            op2.Solver({'linear_solver': 'gmres', 'maximum_iterations': 25},
                       linear_solver='cgs')
