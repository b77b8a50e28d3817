"""wannier90 output -> the TB file the BSE path reads (utils/param_gen.f90).

The reference's param_gen asks for the calculation type (1 NP, 2 SP, 3 SOC), the Fermi
level and the scissor, then writes tb_hr.dat as: the type, the scissor, E_F, the
lines between "begin unit_cell_cart" and "end unit_cell_cart" of wannier90.win
(wannier90.up.win for SP), and wannier90_hr.dat verbatim (for SP the spin-up and
then the spin-down file).  This module does the same without the shell (awk, cat):

    python -m wantibexos_amd.param_gen          # same prompts, in the current folder

or ``param_gen(folder, kind, efermi, scs)`` from Python.
"""
from __future__ import annotations

import sys
from pathlib import Path

KINDS = {1: "NP", 2: "SP", 3: "SOC"}


def unit_cell_lines(win_text: str) -> list[str]:
    """The lines strictly between 'begin unit_cell_cart' and 'end unit_cell_cart' (the
    reference's awk filter: the begin line switches printing on for the next line, the
    end line switches it off before it is printed)."""
    out, on = [], False
    for line in win_text.splitlines():
        if "begin unit_cell_cart" in line:
            on = True
            continue
        if "end unit_cell_cart" in line:
            on = False
        if on:
            out.append(line)
    return out


def param_gen(folder, kind: int, efermi: float, scs: float, out: str = "tb_hr.dat") -> Path:
    """Write `out` in `folder` from that folder's wannier90 files (see the module text)."""
    if kind not in KINDS:
        raise ValueError("wrong value of DFT calculation type")
    folder = Path(folder)
    if kind == 2:
        win = folder / "wannier90.up.win"
        hrs = [folder / "wannier90.up_hr.dat", folder / "wannier90.dn_hr.dat"]
    else:
        win = folder / "wannier90.win"
        hrs = [folder / "wannier90_hr.dat"]
    lines = [f" {KINDS[kind]}", f" {float(scs)!r}  !scissors operator", f" {float(efermi)!r}   !efermi"]
    lines += unit_cell_lines(win.read_text())
    text = "\n".join(lines) + "\n"
    for h in hrs:
        t = h.read_text()
        text += t if t.endswith("\n") else t + "\n"
    path = folder / out
    path.write_text(text)
    return path


def main() -> int:
    print("Write type of DFT calculation")
    print("1 - Non-Polarized")
    print("2 - Spin-Polarized")
    print("3 - Spin Orbit Coupling")
    kind = int(sys.stdin.readline().split()[0])
    if kind not in KINDS:
        print("wrong value of DFT calculation type")
        return 1
    print("Write Fermi level Energy (eV)")
    efermi = float(sys.stdin.readline().split()[0])
    print("Write Scissors Operator")
    scs = float(sys.stdin.readline().split()[0])
    param_gen(".", kind, efermi, scs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
