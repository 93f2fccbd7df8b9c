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


DISCRETE_FORTRAN = """program disc
  implicit none
  integer, parameter :: dl = kind(1.d0), namelength = 12
  integer :: nfld, n, seed, nlog, ns, k, rep, nrep
  real(dl) :: len, dt, mpl, m2phi, m2psi, g2
  real(dl) :: v(64)
  open(10, file='in.txt')
  read(10,*) nfld, n, len, dt, mpl, seed, m2phi, m2psi, g2, nrep
  open(11, file='LOG.out')
  open(12, file='PSD')
  call head(11, [character(len=namelength) :: "t    ", "a    ", "H    ", "<rho>", "<P>  ", "<lnrho>", "<EOS>", &
                 "<E_KIN>", "E_POT>", "<E_grad>", "GRAVEN", "Ham"])
  call head(12, [character(len=namelength) :: "t  ", "a  ", "H   ", "k   ", "logrho_ave", "dlnrho", "lnr_dlnr", &
                 "lnr_dlnr"])
  do rep = 1, nrep
     read(10,*) nlog
     read(10,*) v(1:nlog)
     write(11,'(32g25.16e2)') v(1:nlog)
     read(10,*) ns
     do k = 1, ns
        read(10,*) v(1:8)
        write(12,'(32g25.16e2)') v(1:8)
     enddo
     write(12,'(a)') "",""
  enddo
contains
  subroutine head(fd, vars)
    integer fd
    character(*) :: vars(:)
    character(512) :: buffer
    integer l
    write (fd,'(a,4(i0,a),3(g32.15e2,a),i0)') "# This is DEFROST revision ", 1, ".", 0, " (", nfld, " fields, ", &
         n, "^3 grid) L=", len, " dt = ", dt, " mpl = ", mpl, " seed = ", seed
    write (fd,'(a,i0,a,i0)') "# MPI rank: ", 0, "; total nodes: ", 1
    write (fd,'(a,i0,a,i0,a)') "# Using ", 1, " y division(s) and ", 1, " z division(s)"
    write (fd,'(a,i0,a,i0)') "# Local y division: ", 0, "; local z division: ", 0
    write (fd,'(a,i0,a,i0)') "# neighbor ranks: ", 0, " <- y -> ", 0
    write (fd,'(a,i0,a,i0)') "# neighbor ranks: ", 0, " <- z -> ", 0
    write (fd,'(a,3(f0.5,a))') "# V(phi,psi) = ", m2phi, "phi^2/2 + ", m2psi, "psi^2/2 + ", g2, "phi^2*phi^2/2"
    write (buffer,'(a,g12.12",",32(g24.12","))') "OUTPUT:", adjustr(vars)
    l = index(buffer, ',', .true.)
    write (fd,'(2a)') "# ", repeat('=', l)
    write (fd,'(2a)') "# ", buffer(1:l-1)//';'
    write (fd,'(2a)') "# ", repeat('=', l)
  end subroutine head
end program disc
"""


def _fortran_discrete_files(workdir, header, reps):
    src = os.path.join(workdir, "disc.f90")
    with open(src, "w") as fh:
        fh.write(DISCRETE_FORTRAN)
    subprocess.run([FC, "-O0", src, "-o", os.path.join(workdir, "disc")], check=True, cwd=workdir)
    with open(os.path.join(workdir, "in.txt"), "w") as fh:
        fh.write(" ".join(repr(x) for x in header) + f" {len(reps)}\n")
        for log, rows in reps:
            fh.write(f"{len(log)}\n" + " ".join(repr(float(x)) for x in log) + "\n")
            fh.write(f"{len(rows)}\n")
            for r in rows:
                fh.write(" ".join(repr(float(x)) for x in r) + "\n")
    subprocess.run([os.path.join(workdir, "disc")], check=True, cwd=workdir)
    out = {}
    for name in ("LOG.out", "PSD"):
        with open(os.path.join(workdir, name), "rb") as fh:
            out[name] = fh.read()
    return out


def _psd_rows(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(line[25 * q:25 * (q + 1)]) for q in range(8)])
    return np.array(rows)


@pytest.mark.parametrize("stencil", [(-6.0, 1.0, 0.0, 0.0, 1.0), (-128.0, 14.0, 3.0, 1.0, 30.0)],
                         ids=["7pt", "27pt"])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_discrete_output_matches_oracle_and_fortran_format(gpu, tmp_path, stencil, precision):
    """Discrete-tree make_output (LOG.out line and PSD block with their head()
    headers): values against the oracle's get_variables / output_reg /
    output_log_file restatement, files byte for byte against the same values
    written by amdflang with the reference's format statements, two calls
    appended to the same files."""
    print("case discrete output", stencil, precision, flush=True)
    if not shutil.which(FC) and not os.path.exists(FC):
        pytest.skip("amdflang not installed")
    n, L, mp = 16, 10.0, (1.0, 1e4, 2.0)
    rng = np.random.default_rng(33)
    fld = np.stack([1.0 + 1e-2 * rng.standard_normal((n, n, n)), 1e-2 * rng.standard_normal((n, n, n))])
    fldp = np.stack([-0.7 + 1e-2 * rng.standard_normal((n, n, n)), 1e-2 * rng.standard_normal((n, n, n))])
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    gdir = tmp_path / "gpu"
    gdir.mkdir()
    reps, vals = [], []
    ns = int(np.sqrt(3.0) * (n // 2) + 2)
    with gpu.Lattice(n, nfld=2, length=L, precision=prec, laplacian=gpu.LAP_STENCIL, stencil=stencil,
                     model_params=mp) as lat:
        lat.set_state(fld, fldp, 1.3, -0.4)
        for t in (0.0, 0.25):
            if t > 0:
                lat.symp4(5e-3, 5)
            v, gpsd = lat.discrete_output(t, 5e-3, 2e5, 7, str(gdir))
            f, p, y, py = lat.get_state()
            sim = O.Simulation(n, 2, L, O.COSMIC, lap_kind=O.LAP_STENCIL, stencil=stencil, model=O.MODEL_TWOFIELD,
                               mp=mp)
            sim.set_state(f, p, y, py)
            olog, _, _, opsd = sim.discrete_output(t)
            vals.append((v, olog, opsd, gpsd))
    rows = _psd_rows(gdir / "PSD")
    assert rows.shape == (2 * ns, 8)
    for r, (v, olog, opsd, gpsd) in enumerate(vals):
        tol = 1e-11 if precision == "f64" else 1e-6
        # lnrhoave and eosave are averages of O(1) numbers with O(1e-4)
        # spread, the rest are sums of positive terms or fluctuation means
        assert np.allclose(v, olog, rtol=tol, atol=tol * 1e-2), (r, v, olog)
        got = rows[r * ns:(r + 1) * ns]
        assert np.allclose(got[:, 3], np.arange(ns) * (2 * np.pi / L), rtol=1e-15)
        ptol = 1e-9 if precision == "f64" else 2e-3
        for ch in range(4):
            ref = opsd[ch]
            assert np.allclose(gpsd[ch], ref, rtol=ptol, atol=ptol * np.abs(ref).max()), (r, ch)
            assert np.allclose(got[:, 4 + ch], gpsd[ch], rtol=1e-15, atol=1e-15 * np.abs(gpsd[ch]).max())
        dk = 2 * np.pi / L
        reps.append((v, [[v[0], v[1], v[2], b * dk] + [gpsd[ch][b] for ch in range(4)] for b in range(ns)]))
    # byte compatibility with the Fortran writer, given the same values
    fdir = tmp_path / "fortran"
    fdir.mkdir()
    want = _fortran_discrete_files(str(fdir), [2, n, L, 5e-3, 2e5, 7, mp[0], mp[2], mp[1]], reps)
    for name in ("LOG.out", "PSD"):
        with open(gdir / name, "rb") as fh:
            have = fh.read()
        assert have == want[name], (name, have[:400], want[name][:400])
