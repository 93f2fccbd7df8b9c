"""make_output / dump_rho writer: energy_spec.out and spectrum.out lines in
the reference's '(30(ES22.15,2X))' format, checked value by value against
lat_energies / lat_output_spectra and byte by byte against the same values
written by a Fortran program with that format (amdflang)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FC = "/opt/rocm/bin/amdflang"

FORTRAN = """program fmt
  implicit none
  integer, parameter :: dl = kind(1.d0)
  real(dl) :: v(64)
  integer :: n, ios
  character(len=4096) :: line
  open(10, file='in.txt')
  open(11, file='out.txt')
  do
     read(10, '(A)', iostat=ios) line
     if (ios /= 0) exit
     read(line, *) n
     read(line, *) n, v(1:n)
     write(11, '(30(ES22.15,2X))') v(1:n)
  end do
end program fmt
"""


def fortran_format(rows, workdir):
    src = os.path.join(workdir, "fmt.f90")
    with open(src, "w") as fh:
        fh.write(FORTRAN)
    subprocess.run([FC, "-O0", src, "-o", os.path.join(workdir, "fmt")], check=True, cwd=workdir)
    with open(os.path.join(workdir, "in.txt"), "w") as fh:
        for r in rows:
            fh.write(f"{len(r)} " + " ".join(repr(float(x)) for x in r) + "\n")
    subprocess.run([os.path.join(workdir, "fmt")], check=True, cwd=workdir)
    with open(os.path.join(workdir, "out.txt")) as fh:
        return fh.read().splitlines()


@pytest.mark.parametrize("tree", [0, 1])
def test_make_output_matches_energies_and_fortran_format(gpu, tmp_path, tree):
    print("case make_output", tree, flush=True)
    n, L = 16, 10.0
    fld = np.empty((2, n, n, n))
    fldp = np.empty_like(fld)
    fld[0] = 1.0 + O.sample(n, L, -0.25, 1.0, 2e3, 3, 0)
    fldp[0] = -0.7 + O.sample(n, L, 0.25, 1.0, 2e3, 3, 1)
    fld[1] = O.sample(n, L, -0.25, 1.0, 2e3, 3, 2)
    fldp[1] = O.sample(n, L, 0.25, 1.0, 2e3, 3, 3)
    with gpu.Lattice(n, nfld=2, length=L, model_params=(1.0, 1e4)) as lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        lat.calc_metric()
        lat.symp4(1e-3, 3)
        e = lat.energies()
        y, yp = lat.scalars
        _, pi_now, _, _ = lat.get_state()
        spec = lat.output_spectra(0)
        lat.make_output(0.003, tmp_path, tree)
        lat.make_output(0.003, tmp_path, tree)   # appends
    energy = (tmp_path / "energy_spec.out").read_text().splitlines()
    assert len(energy) == 2 and energy[0] == energy[1]
    vals = [float(t) for t in energy[0].split()]
    a = e["a"]
    expect = [0.003, a, e["rho"], e["ke"], e["pe"], e["ge"], e["grav"], (e["rho"] + e["grav"]) / e["rho"],
              -yp ** 2 / 12.0 / a ** 4, e["mean1"], e["mean2"] if tree == 0 else pi_now[0].mean()]
    assert len(vals) == 11
    for x, y_ in zip(vals, expect):
        assert abs(x - y_) <= 2e-15 * max(abs(y_), 1e-300) + 1e-300 or abs(x - y_) <= 1e-14 * abs(y_)
    lines = (tmp_path / "spectrum.out").read_text().split("\n")
    ns = lat.ns
    block = lines[:ns]
    assert lines[ns] == ""   # write(97,*) empty record
    for b, ln in enumerate(block):
        v = [float(t) for t in ln.split()]
        assert len(v) == 1 + 2 * 2 + 2
        assert abs(v[0] - b * 2 * np.pi / L) <= 1e-15 * max(1.0, v[0])
        for q in range(4):
            assert abs(v[1 + q] - spec[q][b]) <= 1e-15 * abs(spec[q][b]) + 1e-300
        assert v[5] == 0.0 and v[6] == 0.0
    if not os.path.exists(FC):
        pytest.skip("amdflang not available for the byte comparison")
    ref = fortran_format([[float(t) for t in energy[0].split()]] +
                         [[float(t) for t in ln.split()] for ln in block[:5]], str(tmp_path))
    assert ref[0] == energy[0]
    assert ref[1:] == block[:5]


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_checkpoint_restore_resumes_run(gpu, tmp_path, precision):
    """checkpoint after k steps, restore into a fresh context, continue: same
    trajectory as the uninterrupted run; files in the reference's layout."""
    print("case checkpoint", precision, flush=True)
    n, L, dt = 16, 10.0, 1e-3
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    fld = np.empty((2, n, n, n))
    fldp = np.empty_like(fld)
    fld[0] = 1.0 + O.sample(n, L, -0.25, 1.0, 2e3, 5, 0)
    fldp[0] = -0.7 + O.sample(n, L, 0.25, 1.0, 2e3, 5, 1)
    fld[1] = O.sample(n, L, -0.25, 1.0, 2e3, 5, 2)
    fldp[1] = O.sample(n, L, 0.25, 1.0, 2e3, 5, 3)
    kw = dict(nfld=2, length=L, model_params=(1.0, 1e4), precision=prec)
    with gpu.Lattice(n, **kw) as lat:
        lat.set_state(fld, fldp, 1.0, 0.0)
        lat.calc_metric()
        lat.symp4(dt, 3)
        mid = lat.get_state()
        lat.checkpoint(tmp_path, 0.003, 3)
        lat.symp4(dt, 3)
        full = lat.get_state()
    raw = np.fromfile(tmp_path / "fld_2.raw", dtype=np.float64).reshape(n, n, n)
    assert np.array_equal(raw, mid[0][1])
    rec = (tmp_path / "checkpt.dat").read_text().split()
    assert float(rec[0]) == mid[2] and float(rec[1]) == mid[3] and float(rec[2]) == 0.003 and int(rec[3]) == 3
    with gpu.Lattice(n, **kw) as lat:
        t, frame = lat.restore(tmp_path)
        assert (t, frame) == (0.003, 3)
        lat.symp4(dt, 3)
        again = lat.get_state()
    tol = 1e-13 if precision == "f64" else 1e-6
    assert np.linalg.norm(again[0] - full[0]) <= tol * np.linalg.norm(full[0])
    assert np.linalg.norm(again[1] - full[1]) <= tol * np.linalg.norm(full[1])
