"""wannier90 files -> TB file (utils/param_gen.f90) and back through the TB reader
(hamiltonian_input_read, module_input_read.f90:32-176): NP and SP round trips."""
import numpy as np

from wantibexos_amd import param_gen, tbmodel


def write_w90(model, folder, hr_name, win_name):
    """wannier90_hr.dat (comment, nw, nvec, ndegen 15 per line, R m n Re Im rows) and a
    .win with a unit_cell_cart block among other keys."""
    tb = folder / "tmp_tb.dat"
    tbmodel.write_tb(model, tb)
    lines = tb.read_text().splitlines()
    (folder / hr_name).write_text("\n".join(lines[6:]) + "\n")      # from the comment line on
    cell = "\n".join(" ".join(repr(float(v)) for v in row) for row in model.rlat)
    (folder / win_name).write_text(f"num_wann = {model.nw}\nbegin unit_cell_cart\n{cell}\nend unit_cell_cart\n"
                                   "begin projections\nMo:d\nend projections\n")
    tb.unlink()


def test_non_polarized_round_trip(tmp_path):
    m = tbmodel.hex2d(5, 6, (3, 3, 1))
    write_w90(m, tmp_path, "wannier90_hr.dat", "wannier90.win")
    out = param_gen.param_gen(tmp_path, 1, efermi=-0.4321, scs=0.25)
    r = tbmodel.read_tb(out)
    assert r.systype == "NP" and r.scs == 0.25 and r.efermi == -0.4321
    assert np.array_equal(r.rlat, m.rlat)
    assert np.array_equal(r.rvec, m.rvec) and np.array_equal(r.ndeg, m.ndeg)
    assert np.array_equal(r.hop, m.hop) and np.array_equal(r.ihop, m.ihop)


def test_spin_polarized_blocks(tmp_path):
    up = tbmodel.hex2d(6, 4, (3, 3, 1))
    dn = tbmodel.hex2d(7, 5, (3, 3, 1))
    dn.rlat = up.rlat
    write_w90(up, tmp_path, "wannier90.up_hr.dat", "wannier90.up.win")
    write_w90(dn, tmp_path, "wannier90.dn_hr.dat", "wannier90.dn.win")
    r = tbmodel.read_tb(param_gen.param_gen(tmp_path, 2, efermi=0.1, scs=0.0))
    assert r.systype == "SP" and r.nw == 9 and r.nvec == up.nvec + dn.nvec
    assert np.array_equal(r.hop[:up.nvec, :4, :4], up.hop)
    assert np.array_equal(r.hop[up.nvec:, 4:, 4:], dn.hop)
    assert not r.hop[:up.nvec, 4:, :].any() and not r.hop[up.nvec:, :4, :].any()


def test_unit_cell_filter_matches_awk_ranges():
    win = "a\nbegin unit_cell_cart\n1 0 0\n0 1 0\n0 0 1\nend unit_cell_cart\nb\n"
    assert param_gen.unit_cell_lines(win) == ["1 0 0", "0 1 0", "0 0 1"]
