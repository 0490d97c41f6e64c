"""field.h5 output and restart of the drop-in (write_hdf5_field,
write_tecplot_field, initialize_from_hdf5; output.cpp:19-203,
initialization.cpp:270-420, main.cpp:69-111, maincycle.cpp:94-98).

The files are read back with HDF5's own h5dump / h5diff, a reader
independent of the drop-in:

- layout: the attribute and dataset names, their HDF5 types and the
  (nz, ny, nx) shape the reference's writer declares, the stored values;
- restart reader: read -> write reproduces the file exactly; a missing
  file and one of a smaller grid are refused with the reference's messages;
- on the device: the output schedule of time_advance (files after the step
  that follows an output time, at t = 0 and at TotalTime), faithful and
  device-resident files identical, and a run restarted from a mid-run file
  ends where the uninterrupted run ends.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cases import state

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ns3d_wcns_amd", "lib")
DRIVER = os.path.join(LIB, "ns3d_time_advance")


def h5tool(name):
    path = shutil.which(name) or os.path.join("/opt/conda/bin", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not available")
    return path


def attributes(path):
    """name -> (HDF5 type, value) of the root attributes, by h5dump."""
    out = subprocess.run([h5tool("h5dump"), "-m", "%.17g", "-A", path], capture_output=True, text=True,
                         check=True).stdout
    found = {}
    for m in re.finditer(r'ATTRIBUTE "(\w+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SCALAR\s*DATA \{\s*\(0\): (\S+)',
                         out):
        found[m.group(1)] = (m.group(2), float(m.group(3)))
    return found


def datasets(path):
    """name -> (HDF5 type, shape) of the root datasets, by h5dump -H."""
    out = subprocess.run([h5tool("h5dump"), "-H", path], capture_output=True, text=True, check=True).stdout
    found = {}
    for m in re.finditer(r'DATASET "(\w+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)', out):
        found[m.group(1)] = (m.group(2), tuple(int(x) for x in m.group(3).split(",")))
    return found


def dataset(path, name, tmp_path):
    """The raw little-endian doubles of one dataset, by h5dump -b."""
    raw = tmp_path / f"{name}.raw"
    subprocess.run([h5tool("h5dump"), "-b", "LE", "-d", f"/{name}", "-o", str(raw), path], capture_output=True,
                   check=True)
    return np.fromfile(raw, dtype="<f8")


SOLVER_IN = """gamma = 1.4
Pr = 0.72
Ma = {ma!r}
Re = 100
cfl = 0.5
interpolation = WENO5
riemann_solver = HLLC
global_nx = {nx}
global_ny = {ny}
global_nz = {nz}
Lx = {lx!r}
Ly = {ly!r}
Lz = {lz!r}
x0 = {x0!r}
y0 = -1.5
z0 = 0.25
max_steps = {steps}
monitor_Stepfreq = 1
{extra}
"""


def write_case(tmp_path, n, steps=1, extra="", seed=3, lengths=(5.0, 3.0, 2.5), x0=0.5):
    prim, ma = state(n, 0.3, noise=0.01, seed=seed)
    nx, ny, nz = n
    (tmp_path / "solver.in").write_text(SOLVER_IN.format(ma=ma, nx=nx, ny=ny, nz=nz, steps=steps, extra=extra,
                                                         lx=lengths[0], ly=lengths[1], lz=lengths[2], x0=x0))
    with open(tmp_path / "initial.bin", "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        for a in prim:
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
    return prim, ma


def test_print_params_restart_keys(tmp_path):
    f = tmp_path / "solver.in"
    f.write_text("Restart = On\nrestart_file = Output/Time_5/Field.h5\npost_basicfield = no\n"
                 "monitor_energy = ture\n")
    out = subprocess.run([DRIVER, str(f), "--print-params"], capture_output=True, text=True, check=True).stdout
    got = dict(line.split("=", 1) for line in out.split("\n") if "=" in line)
    assert got["restart"] == "1"
    assert got["restart_file"] == "Output/Time_5/Field.h5"  # the value keeps its case
    assert got["post_basicfield"] == "0"
    assert got["monitor_energy"] == "1"  # the reference also accepts "ture"
    from ns3d_wcns_amd import read_solver_params_from_file

    P, _, _ = read_solver_params_from_file(str(f))
    assert (P.restart, P.restart_file, P.post_basicfield, P.monitor_energy) == (
        True, "Output/Time_5/Field.h5", False, True)
    f.write_text("global_nx = 8\n")
    out = subprocess.run([DRIVER, str(f), "--print-params"], capture_output=True, text=True, check=True).stdout
    got = dict(line.split("=", 1) for line in out.split("\n") if "=" in line)
    assert got["restart"] == "0" and got["restart_file"] == "field.h5" and got["post_basicfield"] == "1"
    assert got["global_ny"] == "16" and got["global_nz"] == "16"  # the reference's grid defaults


def test_field_h5_layout(tmp_path):
    n = (5, 4, 3)
    nx, ny, nz = n
    lengths = (5.0, 3.0, 2.5)
    (rho, u, v, w, p), ma = write_case(tmp_path, n, lengths=lengths)
    out = str(tmp_path / "a.h5")
    subprocess.run([DRIVER, str(tmp_path / "solver.in"), "--convert", str(tmp_path / "initial.bin"), out, "1.25"],
                   check=True, capture_output=True)
    attrs = attributes(out)
    dx, dy, dz = lengths[0] / nx, lengths[1] / ny, lengths[2] / nz
    assert attrs == {"global_nx": ("H5T_STD_I32LE", nx), "global_ny": ("H5T_STD_I32LE", ny),
                     "global_nz": ("H5T_STD_I32LE", nz), "time": ("H5T_IEEE_F64LE", 1.25),
                     "dx": ("H5T_IEEE_F64LE", dx), "dy": ("H5T_IEEE_F64LE", dy), "dz": ("H5T_IEEE_F64LE", dz)}
    names = ("X", "Y", "Z", "rho", "u", "v", "w", "E", "p", "T")
    assert datasets(out) == {k: ("H5T_IEEE_F64LE", (nz, ny, nx)) for k in names}
    got = {k: dataset(out, k, tmp_path).reshape(nz, ny, nx) for k in names}
    for k, a in (("rho", rho), ("u", u), ("v", v), ("w", w), ("p", p)):
        assert np.array_equal(got[k], a), k
    # cell coordinates origin + index * spacing, x fastest
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    assert np.array_equal(got["X"], 0.5 + i3 * dx)
    assert np.array_equal(got["Y"], -1.5 + j3 * dy)
    assert np.array_equal(got["Z"], 0.25 + k3 * dz)
    # E and T from the primitives (primitiveToConserved, field_structures.h:468-488)
    g = 1.4
    rgas = 1.0 / (ma * ma * g)
    assert np.allclose(got["E"], rho * (p / ((g - 1) * rho) + 0.5 * (u * u + v * v + w * w)), rtol=1e-15, atol=0)
    assert np.allclose(got["T"], p / (rho * rgas), rtol=1e-15, atol=0)


def test_restart_reader_round_trip_and_errors(tmp_path):
    n = (6, 5, 4)
    write_case(tmp_path, n)
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    solver = str(tmp_path / "solver.in")
    subprocess.run([DRIVER, solver, "--convert", str(tmp_path / "initial.bin"), a, "0.5"], check=True,
                   capture_output=True)
    subprocess.run([DRIVER, solver, "--convert", a, b, "0.5"], check=True, capture_output=True)
    diff = subprocess.run([h5tool("h5diff"), a, b], capture_output=True, text=True)
    assert diff.returncode == 0, diff.stdout
    # the reader selects the current grid's block: HDF5 accepts the selection
    # and refuses the read of a smaller file
    big = tmp_path / "big"
    big.mkdir()
    write_case(big, (7, 5, 4))
    r = subprocess.run([DRIVER, str(big / "solver.in"), "--convert", a, str(big / "c.h5")], capture_output=True,
                       text=True)
    assert r.returncode == 1
    assert "Failed to read dataset 'rho'" in r.stderr
    r = subprocess.run([DRIVER, solver, "--convert", str(tmp_path / "missing.h5"), b], capture_output=True, text=True)
    assert r.returncode == 1
    assert "Failed to open HDF5 file: " in r.stderr


def read_final(path, n):
    nx, ny, nz = n
    cells = (nx + 6) * (ny + 6) * (nz + 6)
    raw = np.fromfile(path, dtype=np.uint8)
    nsteps = int(np.frombuffer(raw[:4].tobytes(), dtype=np.int32)[0])
    d = np.frombuffer(raw[4:].tobytes(), dtype=np.float64)
    return d[:nsteps], d[nsteps:nsteps + 5 * cells].reshape(5, nz + 6, ny + 6, nx + 6)[:, 3:-3, 3:-3, 3:-3]


@pytest.mark.gpu
def test_output_schedule_and_restart(gpu, tmp_path):
    n = (12, 10, 8)
    nx, ny, nz = n
    dt = 2.0 ** -9  # exact multiples: the output times are hit without rounding
    sched = f"dt_fixed = {dt!r}\noutput_Timefreq = {3 * dt!r}\nTotalTime = {8 * dt!r}\n"
    runs = {}
    for mode in ("faithful", "resident"):
        d = tmp_path / mode
        d.mkdir()
        write_case(d, n, steps=100, extra=sched, lengths=(2 * np.pi,) * 3, x0=0.0)
        args = [DRIVER, "solver.in", "initial.bin", "final.bin", "--write-fields"]
        if mode == "resident":
            args.insert(4, "resident")
        print("running", mode, " ".join(args), flush=True)
        subprocess.run(args, cwd=d, check=True, timeout=120, capture_output=True)
        runs[mode] = d
    # output after the step following each output time (3 dt -> step 4, 6 dt
    # -> step 7), plus t = 0 and TotalTime: directory tag int(t * 1e5)
    a = runs["faithful"]
    tags = sorted(os.listdir(a / "output"))
    assert tags == sorted(f"time_{int(k * dt * 100000)}" for k in (0, 4, 7, 8)), tags
    for tag, k in ((f"time_{int(4 * dt * 100000)}", 4), (f"time_{int(8 * dt * 100000)}", 8)):
        assert attributes(str(a / "output" / tag / "field.h5"))["time"][1] == k * dt
    for tag in tags:
        diff = subprocess.run([h5tool("h5diff"), str(a / "output" / tag / "field.h5"),
                               str(runs["resident"] / "output" / tag / "field.h5")], capture_output=True, text=True)
        assert diff.returncode == 0, (tag, diff.stdout)
    dts, fa = read_final(a / "final.bin", n)
    assert len(dts) == 8 and np.all(dts == dt)
    # the file at TotalTime holds the final state
    last = str(a / "output" / f"time_{int(8 * dt * 100000)}" / "field.h5")
    assert np.array_equal(dataset(last, "rho", tmp_path).reshape(nz, ny, nx), fa[0])
    assert np.array_equal(dataset(last, "E", tmp_path).reshape(nz, ny, nx), fa[4])
    # restart from the file written after step 4 and run the last 4 steps
    b = tmp_path / "restart"
    b.mkdir()
    mid = a / "output" / f"time_{int(4 * dt * 100000)}" / "field.h5"
    write_case(b, n, steps=100, lengths=(2 * np.pi,) * 3, x0=0.0,
               extra=f"dt_fixed = {dt!r}\nTotalTime = {4 * dt!r}\noutput_Timefreq = 1.0\n"
                     f"restart = true\nrestart_file = {mid}\n")
    print("running restart", flush=True)
    subprocess.run([DRIVER, "solver.in", "-", "final.bin"], cwd=b, check=True, timeout=120, capture_output=True)
    dtb, fb = read_final(b / "final.bin", n)
    assert len(dtb) == 4 and np.all(dtb == dt)
    # the restart recomputes E from p (main.cpp:99): ulp-level differences
    for q in range(5):
        err = np.max(np.abs(fb[q] - fa[q])) / np.max(np.abs(fa[q]))
        assert err < 1e-12, (q, err)
