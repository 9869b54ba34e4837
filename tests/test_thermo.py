"""Property tables (heated channel, table mode): the host set-up restatement
(chs_prop_table_setup) and the GPU THERM_PROP_UPDATE_FROM_DH lookup, both bit-exact against
the reference build's own outputs (fixture tests/golden/props_2048.npz)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def fx(golden):
    return golden("props_2048.npz")


def table_from_fixture(fx):
    t = {k[2:]: (int(v) if v.ndim == 0 and k[2:].startswith("I") or k == "t_N" else v)
         for k, v in fx.items() if k.startswith("t_")}
    return t


def test_synthetic_table_is_the_fixture_table(pkg, fx):
    from chapsim_amd import thermo
    np.testing.assert_array_equal(thermo.synthetic_table(), fx["rows"])


def test_table_setup_bit_exact(pkg, fx):
    from chapsim_amd import thermo
    mine = thermo.table_setup(fx["rows"], float(fx["T0"]))
    ref = table_from_fixture(fx)
    for k, v in mine.items():
        np.testing.assert_array_equal(np.asarray(v), np.asarray(ref[k]), err_msg=k)


def test_table_is_monotone_in_rho_h(pkg, fx):
    # the reference refuses tables whose rho*h is not monotone (thermo_properties_table_
    # searching.f90:1015-1022); the workload's table must pass that check
    dh = fx["t_DH"]
    assert np.all(np.diff(dh) > 0)
    h = fx["t_H"]
    assert np.all(np.diff(h) > 0)


@pytest.mark.gpu
def test_gpu_lookup_bit_exact(pkg, fx):
    from chapsim_amd import thermo
    pt = thermo.PropTable(table_from_fixture(fx))
    out = pt.from_dh(fx["dh"])
    for i, name in enumerate(("H", "T", "D", "M", "K", "Cp", "B")):
        np.testing.assert_array_equal(out[i], fx["props"][i], err_msg=name)
    pt.close()


@pytest.mark.gpu
def test_gpu_lookup_out_of_range_is_an_error(pkg, fx):
    from chapsim_amd import thermo
    pt = thermo.PropTable(table_from_fixture(fx))
    hi = fx["t_DH"][-1]
    with pytest.raises(pkg.ChsError, match="outside the property table"):
        pt.from_dh(np.array([0.0, hi + 1e-6]))
    pt.close()
