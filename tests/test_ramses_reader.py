"""Native RAMSES particle reader (amdahop.ramses, ahop_ramses_count / ahop_ramses_read) against
the restatement of _read_ramses_new_1010 (oracle/ramses_reader.py, src/input_output.py:624-760)
on part files written record by record here: both layouts (Ra3: families from birth time and
id sign; Ra4: family byte), families all / dm / star, with and without the stellar payload,
1 and 4 threads, and the reference's errors.  Host code only: runs without a GPU."""
import os

import numpy as np
import pytest

from amdahop import ramses
from oracle import ramses_reader as ORR


write_part = ramses.write_part_file


def snapshot(tmp_path, rver, ncpu=5, ndm=3000, nstar=800, nother=150, seed=0):
    """Dark matter with ids 1..ndm, stars with ids 1..nstar (negative birth epochs), and
    particles of neither family, spread over ncpu files in random order."""
    rng = np.random.default_rng(seed)
    n = ndm + nstar + nother
    pos = rng.uniform(0, 1, (n, 3))
    vel = rng.normal(size=(n, 3)) * 1e-3
    mass = rng.uniform(1e-9, 2e-9, n)
    ids = np.concatenate([rng.permutation(ndm) + 1, rng.permutation(nstar) + 1,
                          -(np.arange(nother) + 1)])
    birth = np.concatenate([np.zeros(ndm), -rng.uniform(0.5, 3.0, nstar), np.zeros(nother)])
    metal = np.concatenate([np.zeros(ndm), rng.uniform(0, 0.03, nstar), np.zeros(nother)])
    fam = np.concatenate([np.ones(ndm), np.full(nstar, 2), np.zeros(nother)]).astype(np.int8)
    order = rng.permutation(n)
    d = tmp_path / "output_00042"
    d.mkdir()
    with open(d / "info_00042.txt", "w") as f:
        f.write("ncpu        =          5\nscale_l     =  0.3085677581467192E+27\n"
                "scale_t     =  0.4706028200261398E+18\n")
    for icpu, part in enumerate(np.array_split(order, ncpu), start=1):
        write_part(str(d / f"part_00042.out{icpu:05d}"), ncpu, pos[part], vel[part], mass[part],
                   ids[part], rver, birth[part], metal[part], fam[part])
    snapshot.data = dict(pos=pos, vel=vel, mass=mass, ids=ids, birth=birth, metal=metal)
    return str(d)


SCALE_L, SCALE_T = 0.3085677581467192e27, 0.4706028200261398e18


def same(a, b):
    for k in ("ndm", "nstar", "npart"):
        assert a[k] == b[k], k
    for k in ("pos", "vel", "mass", "source_pid", "birth", "metal"):
        if b[k] is None:
            assert a[k] is None, k
        else:
            assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("rver", ["Ra3", "Ra4"])
@pytest.mark.parametrize("family", ["all", "dm", "star"])
def test_native_reader_matches_the_restatement(tmp_path, rver, family):
    rep = snapshot(tmp_path, rver)
    for stellar in (True, False):
        ref = ORR.read_particles(rep, "00042", 5, rver, family, stellar, SCALE_L, SCALE_T)
        for nt in (1, 4):
            got = ramses.read_particles(rep, rver, family, stellar, nthreads=nt)
            same(got, ref)
    n = {"all": 3800, "dm": 3000, "star": 800}[family]
    assert got["npart"] == n
    assert np.all(got["pos"] >= -0.5) and np.all(got["pos"] < 0.5)


def test_dense_layout_puts_stars_after_the_dark_matter(tmp_path):
    rep = snapshot(tmp_path, "Ra3", ndm=40, nstar=10, nother=3)
    got = ramses.read_particles(rep, "Ra3", "all")
    dm = ramses.read_particles(rep, "Ra3", "dm")
    st = ramses.read_particles(rep, "Ra3", "star")
    assert np.array_equal(got["pos"][:40], dm["pos"]) and np.array_equal(got["pos"][40:], st["pos"])
    assert np.array_equal(dm["source_pid"], np.arange(1, 41))
    assert np.all(st["birth"] < 0) and np.all(np.isfinite(st["metal"]))
    # each particle where its id puts it, shifted and scaled as _read_ramses_new_1010 does
    w = snapshot.data
    for k, lo, hi in ((0, 0, 40), (40, 40, 50)):
        slot = w["ids"][lo:hi] - 1 + k
        assert np.array_equal(got["pos"][slot], w["pos"][lo:hi] - 0.5)
        assert np.array_equal(got["vel"][slot], w["vel"][lo:hi] * SCALE_L / SCALE_T * 1e-5)
        assert np.array_equal(got["mass"][slot], w["mass"][lo:hi])
    assert np.array_equal(st["birth"][w["ids"][40:50] - 1], w["birth"][40:50])
    assert np.array_equal(st["metal"][w["ids"][40:50] - 1], w["metal"][40:50])


def test_reference_errors(tmp_path):
    rep = snapshot(tmp_path, "Ra4", ncpu=2, ndm=20, nstar=5, nother=0)
    path = os.path.join(rep, "part_00042.out00001")
    ok = open(path, "rb").read()
    # a dark-matter id repeated within the file
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 1, (4, 3))
    write_part(path, 2, pos, pos, np.ones(4), np.array([1, 2, 2, 3]), "Ra4",
               family=np.ones(4))
    with pytest.raises(RuntimeError, match="duplicate|mismatch|Non-dense"):
        ramses.read_particles(rep, "Ra4", "dm")
    # an id beyond the selected count
    write_part(path, 2, pos, pos, np.ones(4), np.array([1, 2, 3, 999]), "Ra4", family=np.ones(4))
    with pytest.raises(RuntimeError, match="outside"):
        ramses.read_particles(rep, "Ra4", "dm")
    # a damaged record marker
    bad = bytearray(ok)
    bad[-1] ^= 0xFF
    open(path, "wb").write(bytes(bad))
    with pytest.raises(RuntimeError, match="marker|past the end"):
        ramses.read_particles(rep, "Ra4", "all")


def ra4_with_descriptor(tmp_path, fields, ncpu=3, ndm=500, nstar=200, seed=3, name="output_00042"):
    """A Ra4 output whose stellar records follow ``fields`` (the descriptor's names after the
    tag), each record holding a distinct value per particle so that a field read from the
    wrong record is seen."""
    rng = np.random.default_rng(seed)
    n = ndm + nstar
    pos = rng.uniform(0, 1, (n, 3))
    vel = rng.normal(size=(n, 3)) * 1e-3
    mass = rng.uniform(1e-9, 2e-9, n)
    ids = np.concatenate([rng.permutation(ndm) + 1, rng.permutation(nstar) + 1])
    fam = np.concatenate([np.ones(ndm), np.full(nstar, 2)]).astype(np.int8)
    values = {f: rng.uniform(-3.0, 3.0, n) * (k + 1) for k, f in enumerate(fields)}
    if "birth_time" in values:
        values["birth_time"][ndm:] = -np.abs(values["birth_time"][ndm:]) - 0.1
    d = tmp_path / name
    d.mkdir()
    ramses.write_descriptor(str(d), fields)
    order = rng.permutation(n)
    for icpu, part in enumerate(np.array_split(order, ncpu), start=1):
        recs = [values[f][part] for f in fields]
        write_part(str(d / f"part_{name[-5:]}.out{icpu:05d}"), ncpu, pos[part], vel[part],
                   mass[part], ids[part], "Ra4", np.where(fam[part] == 2, -1.0, 0.0),
                   None, fam[part], stellar_records=recs)
    star_ids = ids[ndm:]
    expect = {f: np.empty(nstar) for f in fields}
    for f in fields:
        expect[f][star_ids - 1] = values[f][ndm:]
    return str(d), expect


DESCRIPTOR_ORDERS = [
    ["birth_time", "metallicity"],
    ["metallicity", "birth_time"],
    ["birth_time", "iord", "metallicity", "initial_mass", "chem_Fe", "chem_O", "ptype"],
    ["metallicity", "chem_H", "birth_time", "chem_D", "initial_mass"],
]


@pytest.mark.parametrize("fields", DESCRIPTOR_ORDERS)
def test_ra4_descriptor_orders(tmp_path, fields):
    """Stellar records read where part_file_descriptor.txt places them (_ra4_stellar_descriptor
    and _read_ra4_stellar_record_block, src/input_output.py:547-620): birth times,
    metallicities, initial masses and chem abundances, other fields skipped; native reader and
    restatement against the values written."""
    rep, expect = ra4_with_descriptor(tmp_path, fields)
    present, has_m0, elems = ramses.descriptor(rep)
    assert present and has_m0 == ("initial_mass" in fields)
    assert elems == tuple(e for e in ramses.CHEM_ELEMENTS if f"chem_{e}" in fields)
    for family in ("all", "star"):
        got = ramses.read_particles(rep, "Ra4", family, nthreads=2)
        ref = ORR.read_particles(rep, "00042", 3, "Ra4", family, True)
        same(got, ref)
        for k in ("m0", "chem"):
            if ref[k] is None:
                assert got[k] is None
            else:
                assert np.array_equal(got[k], ref[k], equal_nan=True), k
        assert np.array_equal(got["birth"], expect["birth_time"])
        assert np.array_equal(got["metal"], expect["metallicity"])
        if "initial_mass" in fields:
            assert np.array_equal(got["m0"], expect["initial_mass"])
        for i, e in enumerate(ramses.CHEM_ELEMENTS):
            if f"chem_{e}" in fields:
                assert np.array_equal(got["chem"][:, i], expect[f"chem_{e}"])
            elif got["chem"] is not None:
                assert np.all(np.isnan(got["chem"][:, i]))
    # without the stellar payload nothing after the tag is read
    assert ramses.read_particles(rep, "Ra4", "all", stellar=False)["birth"] is None


@pytest.mark.parametrize("fields,msg", [
    (["birth_time", "initial_mass"], "initial_mass but no metallicity"),
    (["birth_time", "initial_mass", "metallicity"], "Unsupported Ra4 initial_mass record order"),
])
def test_ra4_descriptor_errors(tmp_path, fields, msg):
    rep, _ = ra4_with_descriptor(tmp_path, fields)
    with pytest.raises(RuntimeError, match=msg):
        ramses.descriptor(rep)
    with pytest.raises(RuntimeError, match=msg):
        ramses.read_particles(rep, "Ra4", "all")
    with pytest.raises(RuntimeError, match=msg):
        ORR.read_particles(rep, "00042", 3, "Ra4", "all", True)


def test_ra4_descriptor_without_birth_time_is_an_error(tmp_path):
    """Stars need a birth time and a metallicity (the reference fails on them otherwise)."""
    rep, _ = ra4_with_descriptor(tmp_path, ["metallicity", "chem_O"])
    with pytest.raises(RuntimeError, match="birth_time or metallicity"):
        ramses.read_particles(rep, "Ra4", "all")
