"""Host-side API of ``pyop2_amd.op2`` (CPU): construction, validation errors
and caching behave as in the reference (``test/unit/test_api.py``,
``test_caching.py:491-548``, ``test_indirect_loop.py:109-129``,
``test_matrices.py:575-580``), and a par_loop without a device fails loudly."""
import numpy as np
import pytest

from pyop2_amd import exceptions as ex, op2
from pyop2_amd.configuration import configuration


def test_set_sizes_and_parts():
    s = op2.Set((3, 7, 9), "s")
    assert (s.core_size, s.size, s.total_size) == (3, 7, 9)
    assert (s.core_part.offset, s.core_part.size) == (0, 3)
    assert (s.owned_part.offset, s.owned_part.size) == (3, 4)
    assert op2.Set(5).sizes == (5, 5, 5)
    with pytest.raises(ex.SizeTypeError):
        op2.Set("five")
    with pytest.raises(ex.NameTypeError):
        op2.Set(5, 12)


def test_subset():
    s = op2.Set(10)
    ss = s([7, 2, 2, 5])
    assert list(ss.indices) == [2, 5, 7]
    assert ss.superset is s and ss.total_size == 3
    assert list(ss([0, 2]).indices) == [2, 7]
    with pytest.raises(ex.SubsetIndexOutOfBounds):
        op2.Subset(s, [11])


def test_set_operations():
    """test_subset.py:261-330: union/intersection/difference/symmetric
    difference between Sets and Subsets."""
    a = op2.Set(10)
    assert a.union(a) is a and a.intersection(a) is a
    assert a.difference(a)._indices.size == 0 and a.symmetric_difference(a)._indices.size == 0
    b = op2.Subset(a, np.array([2, 3, 5, 7], dtype=np.int32))
    assert a.union(b) is a and a.intersection(b) is b
    assert list(a.difference(b)._indices) == [0, 1, 4, 6, 8, 9]
    assert list(a.symmetric_difference(b)._indices) == [0, 1, 4, 6, 8, 9]
    assert b.union(a) is a and b.intersection(a) is b
    assert b.difference(a)._indices.size == 0
    assert list(b.symmetric_difference(a)._indices) == [0, 1, 4, 6, 8, 9]
    c = op2.Subset(a, np.array([2, 4, 6, 8], dtype=np.int32))
    assert list(b.union(c)._indices) == [2, 3, 4, 5, 6, 7, 8]
    assert list(b.intersection(c)._indices) == [2]
    assert list(b.difference(c)._indices) == [3, 5, 7]
    assert list(b.symmetric_difference(c)._indices) == [3, 4, 5, 6, 7, 8]


def test_permuted_composed_map_api():
    """map.py:171-287: a PermutedMap keeps its base map's values and checks
    the permutation; a ComposedMap checks its chain and stores no values."""
    s1, s2, s3 = op2.Set(3), op2.Set(4), op2.Set(5)
    m = op2.Map(s2, s3, 2, values=[[0, 1], [1, 2], [2, 3], [3, 4]])
    pm = op2.PermutedMap(m, [1, 0])
    assert (pm.values == m.values).all() and list(pm.permutation) == [1, 0]
    assert (pm.iterset, pm.toset, pm.arity) == (s2, s3, 2)
    with pytest.raises(ValueError):
        op2.PermutedMap(m, [0, 0])
    c = op2.ComposedMap(m, op2.Map(s1, s2, 1, values=[3, -1, 0]))
    assert c.flattened_maps == (m, c.maps_[1])
    assert c._table.tolist() == [[3, 4], [-1, -1], [0, 1]]      # a -1 anywhere masks the entry
    with pytest.raises(RuntimeError):
        c.values_with_halo
    with pytest.raises(ex.MapTypeError):
        op2.ComposedMap(m, op2.Map(s1, s3, 1, values=[0, 1, 2]))
    with pytest.raises(NotImplementedError):
        op2.PermutedMap(c, [1, 0])


def test_mixed_types_api():
    """types/dataset.py:166-312, map.py:297-437, dat.py:646-880, mat.py:86-114:
    components, sizes, block structure and the errors of mixed objects."""
    mset = op2.MixedSet((op2.Set(3), op2.Set(4)))
    elem, node = mset
    mds = mset ** 2
    assert [d.cdim for d in mds] == [2, 2] and mds.cdim == 4 and mds.set is mset
    assert op2.MixedDataSet(mset, (1, 3)).dim == ((1,), (3,))
    mmap = op2.MixedMap((op2.Map(elem, elem, 1, [0, 1, 2]), op2.Map(elem, node, 2, [0, 1, 1, 2, 2, 3])))
    assert mmap.arity == 3 and mmap.arities == (1, 2) and mmap.arange == (0, 1, 3)
    assert mmap.values.tolist() == [[0, 0, 1], [1, 1, 2], [2, 2, 3]]
    assert mmap.iterset is elem and mmap.toset == mset
    with pytest.raises(ex.MapTypeError):
        op2.MixedMap((op2.Map(elem, node, 1, [0, 1, 2]), op2.Map(node, node, 1, [0, 1, 2, 3])))
    md = op2.MixedDat(mset ** 2)
    assert [d.data_ro.shape for d in md] == [(3, 2), (4, 2)] and len(md) == 2
    with pytest.raises(ex.MapValueError):      # component tosets differ from the Dat's sets
        md(op2.READ, op2.MixedMap((mmap[1], mmap[0])))
    sp = op2.Sparsity((mset ** 1, mset ** 1), {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap)
                                               for j, cm in enumerate(mmap)})
    assert sp.shape == (2, 2) and sp.nested and sp[1, 0].dsets == (node ** 1, elem ** 1)
    A = op2.Mat(sp)
    assert type(A).__name__ == "MixedMat" and A[0, 1].sparsity is sp[0, 1] and len(list(A)) == 4
    with pytest.raises(RuntimeError):
        op2.Sparsity((mset ** 1, mset ** 1), {(2, 0): [(mmap[0], mmap[0], None)]})


def test_dataset_identity_cache():
    s = op2.Set(4)
    assert s ** 2 is s ** 2
    assert s ** 2 is not s ** 3
    assert (s ** (2, 3)).cdim == 6
    with pytest.raises(ex.DimTypeError):
        op2.DataSet(s, "x")


def test_map_validation():
    a, b = op2.Set(3), op2.Set(4)
    m = op2.Map(a, b, 2, [0, 1, 2, 3, 1, 0])
    assert m.values.shape == (3, 2) and m.values.dtype == np.int32
    with pytest.raises(ex.DataValueError):
        op2.Map(a, b, 2, [0, 1])
    with pytest.raises(ex.SetTypeError):
        op2.Map("a", b, 2)
    with pytest.raises(ex.ArityTypeError):
        op2.Map(a, b, 2.5)


def test_dat_host_semantics(tmp_path):
    s = op2.Set(5)
    d = op2.Dat(s ** 2, np.arange(10.0).reshape(5, 2))
    assert d.data.shape == (5, 2) and d.dtype == np.float64
    ro = d.data_ro
    with pytest.raises(ValueError):
        ro[0, 0] = 1.0
    d.data[0, 0] = 42.0
    assert d.data_ro[0, 0] == 42.0
    z = op2.Dat(s)
    assert (z.data_ro == 0).all() and z.dtype == np.float64
    d.save(str(tmp_path / "d"))
    e = op2.Dat(s ** 2)
    e.load(str(tmp_path / "d"))
    assert np.array_equal(e.data_ro, d.data_ro)
    d.zero()
    assert (d.data_ro == 0).all()
    c = op2.Dat(e)
    assert np.array_equal(c.data_ro, e.data_ro)
    with pytest.raises(ex.DataSetTypeError):
        op2.Dat("s")


def test_dat_call_checks_map_target():
    a, b, c = op2.Set(3), op2.Set(4), op2.Set(4)
    m = op2.Map(a, b, 1, [0, 1, 2])
    d = op2.Dat(c)
    with pytest.raises(ex.MapValueError):
        d(op2.READ, m)
    with pytest.raises(ex.ModeValueError):
        d("bogus", m)


def test_global():
    g = op2.Global(2, [1.0, 2.0], name="g")
    assert g.shape == (2,) and list(g.data_ro) == [1.0, 2.0]
    g.data = [3.0, 4.0]
    assert list(g.data_ro) == [3.0, 4.0]
    with pytest.raises(ex.ModeValueError):
        g(op2.WRITE)
    with pytest.raises(ex.ModeValueError):
        op2.Constant(1, 1.0)(op2.INC)


def test_parloop_validation_errors():
    it, ind = op2.Set(4), op2.Set(4)
    x = op2.Dat(ind, np.zeros(4))
    k = op2.Kernel("static void w(double *x) { *x = 1.0; }", "w")
    with pytest.raises(ex.MapValueError):     # iterset of the map differs
        op2.par_loop(k, it, x(op2.WRITE, op2.Map(op2.Set(4), ind, 1, [0, 1, 2, 3])))
    with pytest.raises(ex.MapValueError):     # toset differs from the Dat's set
        op2.par_loop(k, it, x(op2.WRITE, op2.Map(it, op2.Set(4), 1, [0, 1, 2, 3])))
    with pytest.raises(ex.MapValueError):     # uninitialised map
        op2.par_loop(k, it, x(op2.WRITE, op2.Map(it, ind, 1)))
    with pytest.raises(ex.SetTypeError):
        op2.par_loop(k, "it", x(op2.WRITE, op2.Map(it, ind, 1, [0, 1, 2, 3])))
    with pytest.raises(ex.KernelTypeError):
        op2.par_loop("not a kernel", it)
    with pytest.raises(ex.MapValueError):     # direct arg on a different set
        op2.par_loop(k, it, op2.Dat(op2.Set(4))(op2.WRITE))


def test_mat_modes_and_sparsity_cache():
    nodes, els = op2.Set(4), op2.Set(2)
    en = op2.Map(els, nodes, 3, [0, 1, 3, 2, 3, 1])
    sp = op2.Sparsity((nodes ** 1, nodes ** 1), [(en, en, None)])
    assert sp is op2.Sparsity((nodes ** 1, nodes ** 1), [(en, en, None)])
    assert sp is not op2.Sparsity((nodes ** 2, nodes ** 2), [(en, en, None)])
    A = op2.Mat(sp)
    for mode in (op2.READ, op2.RW, op2.MAX, op2.MIN):
        with pytest.raises(ex.ModeValueError):
            A(mode, (en, en))
    other = op2.Map(els, nodes, 3, [0, 1, 3, 2, 3, 1])
    with pytest.raises(ex.MapValueError):
        A(op2.INC, (other, other))
    with pytest.raises(ex.MapValueError):
        op2.Sparsity((nodes ** 1, nodes ** 1), [(op2.Map(els, nodes, 1), op2.Map(els, nodes, 1), None)])
    with pytest.raises(RuntimeError):
        op2.Mat(sp, np.int32)


def test_kernel_cache_keys():
    k1 = op2.Kernel("void a(double *x) {}", "a")
    k2 = op2.Kernel("void a(double *x) {}", "a")
    k3 = op2.Kernel("void a(double *x) { }", "a")
    assert k1.cache_key == k2.cache_key != k3.cache_key
    with pytest.raises(ex.NameTypeError):
        op2.Kernel("", 3)


def test_configuration():
    assert configuration["type_check"] is True
    with pytest.raises(ex.ConfigurationError):
        configuration["type_check"] = "yes"
    with pytest.raises(ex.ConfigurationError):
        configuration["strategy"] = "magic"
    op2.init(strategy="coloured")
    assert configuration["strategy"] == "coloured"


def test_parloop_without_device_fails_loudly():
    from pyop2_amd import device as dev
    if dev.torch().cuda.is_available():
        pytest.skip("a device is present")
    it, ind = op2.Set(4), op2.Set(4)
    x = op2.Dat(ind, np.zeros(4))
    m = op2.Map(it, ind, 1, [0, 1, 2, 3])
    with pytest.raises(ex.DeviceError):
        op2.par_loop(op2.Kernel("static void w(double *x) { *x = 1.0; }", "w"), it, x(op2.WRITE, m))


def test_empty_map_and_iterset():
    """test_api.py:1593-1601: over a zero-sized iteration set a map without
    values is not an error, and nothing is launched."""
    s1 = op2.Set(0)
    s2 = op2.Set(10)
    m = op2.Map(s1, s2, 3)
    d = op2.Dat(s2 ** 1, [0] * 10, dtype=int)
    k = op2.Kernel("static void k(int *x) {}", "k")
    op2.par_loop(k, s1, d(op2.READ, m))


def test_iteration_region_and_periodic_api():
    """Iteration regions are the reference's IntEnum (types/__init__.py:14-31);
    a Sparsity stores each pair's regions sorted (ALL when None is given,
    mat.py:139-156); an ExtrudedSet may be periodic and a Map carries its
    offset quotients into the kernel argument."""
    assert [int(r) for r in (op2.ON_BOTTOM, op2.ON_TOP, op2.ON_INTERIOR_FACETS, op2.ALL)] == [1, 2, 3, 4]
    base = op2.Set(3)
    ext = op2.ExtrudedSet(base, 4, extruded_periodic=True)
    assert ext._extruded_periodic and ext.constant_layers and ext.layers == 4
    t = op2.Set(20)
    m = op2.Map(ext, t, 2, [[0, 1], [5, 6], [10, 11]], offset=[1, 1], offset_quotient=[0, 1])
    assert m._global_kernel_arg.offset == (1, 1) and m._global_kernel_arg.offset_quotient == (0, 1)
    sp = op2.Sparsity((t ** 1, t ** 1), [(m, m, (op2.ON_TOP, op2.ON_BOTTOM))])
    assert sp.iteration_regions[(0, 0)] == [(op2.ON_BOTTOM, op2.ON_TOP)]
    sp_all = op2.Sparsity((t ** 1, t ** 1), [(m, m, None)])
    assert sp_all.iteration_regions[(0, 0)] == [(op2.ALL,)]
    pm = op2.PermutedMap(op2.Map(ext, t, 2, [[0, 1], [5, 6], [10, 11]], offset=[1, 2]), [1, 0])
    assert list(pm.offset) == [2, 1]
    with pytest.raises(ValueError):
        op2.Sparsity((t ** 1, t ** 1), [(m, m, (7,))])


def test_global_arithmetic():
    """Globals combine with scalars and Globals from either side
    (glob.py:147-190): the result is a new Global; unary plus copies."""
    g = op2.Global(2, [2.0, 4.0], np.float64, "g")
    h = op2.Global(2, [1.0, 1.0], np.float64, "h")
    assert list((g + h).data_ro) == [3.0, 5.0]
    assert list((1.0 + g).data_ro) == [3.0, 5.0]
    assert list((1.0 - g).data_ro) == [-1.0, -3.0]
    assert list((3.0 * g).data_ro) == [6.0, 12.0]
    assert list((g / 2.0).data_ro) == [1.0, 2.0]
    p = +g
    assert p is not g and list(p.data_ro) == [2.0, 4.0]
    g -= h
    assert list(g.data_ro) == [1.0, 3.0]


def test_large_loops_take_the_64bit_path():
    """Loops whose element x arity or node x 3 products pass 2**31 leave the
    hand-written (int32-indexed) entry points for the generated loops."""
    from types import SimpleNamespace
    from pyop2_amd.registry import fits_32bit_indexing
    small, big, nodes = op2.Set(1000), op2.Set(300_000_000), op2.Set(1000)
    for it, want in ((small, True), (big, False)):
        m = op2.Map(it, nodes, 10)
        assert fits_32bit_indexing(SimpleNamespace(arguments=[SimpleNamespace(maps=(m,))])) is want
    huge_nodes = op2.Set(800_000_000)
    m = op2.Map(small, huge_nodes, 4)
    assert not fits_32bit_indexing(SimpleNamespace(arguments=[SimpleNamespace(maps=(m,))]))


@pytest.fixture
def vdat():
    return op2.Dat(op2.Set(5) ** 2, np.zeros((5, 2)), np.float64, "vdat")


def test_dat_view_assign(vdat):
    """test_dats.py:269-282: writing a view writes the parent's component
    and invalidates the halo."""
    vdat.data[:, 0] = 3
    vdat.data[:, 1] = 4
    comp = op2.DatView(vdat, 1)
    comp.data[:] = 7
    assert not vdat.halo_valid and not comp.halo_valid
    assert all(comp.data == 7) and all(vdat.data[:, 0] == 3) and all(vdat.data[:, 1] == 7)


def test_dat_view_zero(vdat):
    """test_dats.py:284-297: zeroing a view zeroes the component and leaves
    the halo valid."""
    vdat.data[:, 0] = 3
    vdat.data[:, 1] = 4
    comp = op2.DatView(vdat, 1)
    comp.zero()
    assert vdat.halo_valid and comp.halo_valid
    assert all(comp.data == 0) and all(vdat.data[:, 0] == 3)


def test_dat_view_halo_valid(vdat):
    """test_dats.py:299-316: reading a view keeps the version; a writable
    view with halos bumps the parent's version and dirties its halo."""
    comp = op2.DatView(vdat, 1)
    assert vdat.halo_valid and comp.halo_valid
    assert vdat.dat_version == 0 and comp.dat_version == 0
    comp.data_ro_with_halos
    assert vdat.halo_valid and vdat.dat_version == 0 and comp.dat_version == 0
    comp.data_with_halos
    assert not vdat.halo_valid and not comp.halo_valid
    assert vdat.dat_version == 1 and comp.dat_version == 1


def test_dat_view_index_checks(vdat):
    with pytest.raises(op2.IndexValueError):
        op2.DatView(vdat, 2)


def test_dat_version():
    """test_dats.py:143-171: data access, zero and copy each count once."""
    s = op2.Set(5)
    d1 = op2.Dat(s, list(range(5)), dtype=np.float64)
    d2 = op2.Dat(s)
    assert d1.dat_version == 0 and d2.dat_version == 0
    d1.data
    assert d1.dat_version == 1 and d2.dat_version == 0
    d2.data[:] += 1
    assert d1.dat_version == 1 and d2.dat_version == 1
    d1.zero()
    assert d1.dat_version == 2 and d2.dat_version == 1
    d2.copy(d1)
    assert d1.dat_version == 3 and d2.dat_version == 1


def test_mixed_dat_version():
    """test_dats.py:206-232: a MixedDat's version is the sum of its
    components'; zeroing [d1, d1] zeroes d1 twice."""
    s = op2.Set(5)
    d1 = op2.Dat(s, list(range(5)), dtype=np.float64)
    mdat = op2.MixedDat([d1, d1])
    d2 = op2.Dat(s)
    mdat2 = op2.MixedDat([d1, d2])
    assert mdat.dat_version == 0 and mdat2.dat_version == 0
    mdat2.data
    assert d1.dat_version == 1 and d2.dat_version == 1
    assert mdat.dat_version == 2 and mdat2.dat_version == 2
    mdat.zero()
    assert d1.dat_version == 3 and d2.dat_version == 1
    assert mdat.dat_version == 6 and mdat2.dat_version == 4
    d1.zero()
    assert d1.dat_version == 4 and mdat.dat_version == 8 and mdat2.dat_version == 5
