"""The h5py-free HDF5 files against the reference's own output files.

``pyvibdmc_amd.hdf5.write`` must produce the bytes h5py wrote for the reference's
``SimArchivist.save_h5`` (simulation_utilities/sim_archive.py:13-24): reading a reference
file and writing its datasets back (same names, same order, same modification time) gives
the identical file.  Checked on the committed ``sim_info`` fixture (pyvibdmc/sample_sim_data/
tutorial_water_1_sim_info.hdf5) and, where the reference tree is present, on all 30 of its
sample files (sim_info and wave functions).  ``SimInfo`` mirrors
analysis/extract_sim_info.py:9-133."""
import glob
import os
import struct

import numpy as np
import pytest

from pyvibdmc_amd import hdf5
from tests.conftest import GOLDEN

FIXTURE = os.path.join(GOLDEN, 'ref_tutorial_water_1_sim_info.hdf5')
REF_DATA = '/root/reference/pyvibdmc/sample_sim_data'


def rewrite(fn, tmp_path):
    b = open(fn, 'rb').read()
    d = hdf5.read(fn)
    heap = struct.unpack_from('<Q', b, 680 + 24)[0]
    order = sorted(d, key=lambda k: b.index(k.encode() + b'\0', heap))  # insertion order: heap order
    i = b.find(bytes.fromhex('1200080000000000'))  # modification-time message
    mtime = struct.unpack_from('<I', b, i + 12)[0]
    out = str(tmp_path / 'rewritten.hdf5')
    hdf5.write(out, order, [d[k] for k in order], mtime=mtime)
    return b, open(out, 'rb').read()


def test_rewrite_of_reference_sim_info_is_byte_identical(tmp_path):
    b, o = rewrite(FIXTURE, tmp_path)
    assert o == b


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason='reference sample data not present')
def test_rewrite_of_every_reference_file_is_byte_identical(tmp_path):
    files = sorted(glob.glob(f'{REF_DATA}/*.hdf5')) + sorted(glob.glob(f'{REF_DATA}/wfns/*.hdf5'))
    assert len(files) == 30
    for fn in files:
        b, o = rewrite(fn, tmp_path)
        assert o == b, fn


def test_roundtrip_dtypes_shapes_and_small_data(tmp_path):
    """Float64/32 and int64/32 arrays, large (contiguous at the end of the file) and small
    (the small-data block), in several orders."""
    rs = np.random.RandomState(0)
    vals = {'coords': rs.normal(size=(300, 3, 3)), 'desc_wts': rs.uniform(size=300),
            'parent_wts': rs.uniform(size=7).astype(np.float32), 'n': np.arange(5, dtype=np.int32),
            'big_int': rs.randint(-10 ** 12, 10 ** 12, size=(400, 2), dtype=np.int64)}
    for keys in (list(vals), list(vals)[::-1], ['n', 'coords']):
        fn = str(tmp_path / 'r.hdf5')
        hdf5.write(fn, keys, [vals[k] for k in keys])
        back = hdf5.read(fn)
        assert sorted(back) == sorted(keys)
        for k in keys:
            assert back[k].dtype == vals[k].dtype and back[k].shape == vals[k].shape
            np.testing.assert_array_equal(back[k], vals[k])


def test_siminfo_reads_the_reference_results():
    import pyvibdmc_amd as pv
    s = pv.SimInfo(FIXTURE)
    v = s.get_vref()
    assert v.shape == (5000, 2) and s.get_pop().shape == (5000, 2)
    np.testing.assert_array_equal(v[:, 0], np.arange(5000) * 5.0)
    np.testing.assert_array_equal(s.get_atomic_nums(), [1, 1, 10])  # the reference's table order
    zpe = s.get_zpe(onwards=1000, ret_cm=True)
    assert zpe == pytest.approx(np.mean(v[1000:, 1]) / 4.556335281212229e-6, rel=1e-14)
    # the reference's five tutorial runs average 4634.1 +- 2.2 cm^-1
    assert abs(zpe - 4634.1) < 3 * 2.2
    cm = s.get_vref(ret_cm=True)
    np.testing.assert_allclose(cm[:, 1] * 4.556335281212229e-6, v[:, 1], rtol=1e-14)
    assert s.window_avg(blocks=5).shape == (5, 2)


def test_siminfo_wave_functions_and_training(tmp_path):
    import pyvibdmc_amd as pv
    rs = np.random.RandomState(1)
    os.makedirs(tmp_path / 'wfns')
    hdf5.write(str(tmp_path / 'x_sim_info.hdf5'), ['vref_vs_tau', 'pop_vs_tau', 'atomic_nums', 'atomic_masses'],
               [rs.normal(size=(10, 2)), rs.normal(size=(10, 2)), np.array([1, 1, 10]), np.ones(3)])
    w = {}
    for t in (100, 200):
        w[t] = (rs.normal(size=(50 + t // 10, 3, 3)), rs.uniform(size=50 + t // 10), rs.uniform(size=50 + t // 10))
        hdf5.write(str(tmp_path / 'wfns' / f'x_wfn_{t}ts.hdf5'), ['coords', 'desc_wts', 'parent_wts'], list(w[t]))
    s = pv.SimInfo(str(tmp_path / 'x_sim_info.hdf5'))
    cds, dw, pw = s.get_wfns([100, 200], get_parent_wts=True)
    np.testing.assert_array_equal(cds, np.concatenate([w[100][0], w[200][0]]))
    np.testing.assert_array_equal(dw, np.concatenate([w[100][1], w[200][1]]))
    np.testing.assert_array_equal(pw, np.concatenate([w[100][2], w[200][2]]))
    ca, _ = s.get_wfns(100, ret_ang=True)
    np.testing.assert_allclose(ca * (1 / 0.529177), w[100][0], rtol=1e-14)
    hdf5.write(str(tmp_path / 't.hdf5'), ['coords', 'pots'], [w[100][0], w[100][1]])
    c, p = pv.SimInfo.get_training(str(tmp_path / 't.hdf5'))
    np.testing.assert_array_equal(p, w[100][1])
