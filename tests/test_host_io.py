"""Host file formats (ns3d_cpp_amd/host/ns3d_io.cpp) on the CPU: the Tecplot
field writer and readers (output.cpp:21-87, initialization.cpp:35-190), the
spectrum file (EnergySpectrum.cpp:127-133), the turbulence-statistics
history (ana_isotropic.cpp:300-342) and the residual history
(output.cpp:93-128).

The reference writes through iostreams; the formats test checks that the
snprintf conversions used here print the same bytes as iostreams for a set of
edge values.  The readers are pinned by round trips and by the resampling
rule evaluated independently; the dissipation by a spectrum whose sum is
known in closed form."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ns3d_cpp_amd", "host")


# every test below runs twice: built plainly, and built under AddressSanitizer
# and UndefinedBehaviorSanitizer (a report makes the harness exit non-zero)
BUILDS = {"plain": [], "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                                      "-fno-omit-frame-pointer", "-g"]}
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"}


@pytest.fixture(scope="module", params=list(BUILDS))
def tool(request, tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("io") / "host_io_check")
    subprocess.run(["g++", "-O1", "-std=c++17", *BUILDS[request.param], "-I", HOST,
                    os.path.join(ROOT, "tests", "host_io_check.cpp"), os.path.join(HOST, "ns3d_io.cpp"), "-o", exe],
                   check=True)
    env = {**os.environ, **SAN_ENV}

    def run(*args, cwd=None):
        r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, cwd=cwd, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        return r.stdout
    return run


def pattern(q, i, j, k):
    return (q + 1) * 0.5 + i * 0.125 - j * 0.0625 + k * 2.0 + (-3.75 if q == 1 else 0.0)


def parse_tecplot(path):
    with open(path) as f:
        lines = f.read().splitlines()
    z = next(n for n, line in enumerate(lines) if line.startswith("ZONE"))
    return lines[:z + 1], lines[z + 1:]


def test_snprintf_matches_iostream(tool):
    assert "formats equal" in tool("formats")


def test_paths(tool):
    out = tool("paths", 0.5).split()
    assert out == ["output/time_50000/field_rank0.dat", "output/output_time_0.50000"]
    # the directory name truncates time*1e5 (output.cpp:28): 0.29*1e5 < 29000
    assert tool("paths", 0.29).split()[0] == "output/time_28999/field_rank0.dat"


def test_write_tecplot(tool, tmp_path):
    nx, ny, nz, ng = 5, 3, 2, 2
    path = tmp_path / "f.dat"
    tool("write", path, nx, ny, nz, ng, 0.125)
    head, data = parse_tecplot(path)
    assert head == ['TITLE = "NS3D Field Rank 0  Time=1.25000000e-01"',
                    'VARIABLES = "X" "Y" "Z" "rho" "u" "v" "w" "E" "p" "T"',
                    f'ZONE T="rank_0 time=1.25000000e-01" I={nx} J={ny} K={nz} DATAPACKING=POINT']
    assert len(data) == nx * ny * nz
    n = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                x, y, z = -1.0 + i * 0.25, 0.5 + j * 0.125, 2.0 + k * (1.0 / 3.0)
                # file columns rho u v w E p T = arrays 0 1 2 3 4 5 6
                want = [x, y, z] + [pattern(q, i, j, k) for q in range(7)]
                assert data[n] == " ".join("%.8e" % v for v in want)
                n += 1


def write_file(path, nx, ny, nz, rng):
    vals = rng.normal(size=(nz, ny, nx, 10))
    with open(path, "w") as f:
        f.write('TITLE = "t"\nVARIABLES = "X" "Y" "Z" "rho" "u" "v" "w" "E" "p" "T"\n')
        f.write(f"ZONE I={nx} J={ny} K={nz}\n")
        for v in vals.reshape(-1, 10):
            f.write(" ".join(repr(float(a)) for a in v) + "\n")
    return vals


def load_raw(path, nx, ny, nz, ng):
    s = (nz + 2 * ng, ny + 2 * ng, nx + 2 * ng)
    return np.fromfile(path, dtype=np.float64).reshape(5, *s)


def test_read_tecplot_round_trip(tool, tmp_path):
    nx, ny, nz, ng = 4, 3, 5, 3
    vals = write_file(tmp_path / "in.dat", nx, ny, nz, np.random.default_rng(3))
    tool("read", tmp_path / "in.dat", nx, ny, nz, ng, tmp_path / "out.raw")
    a = load_raw(tmp_path / "out.raw", nx, ny, nz, ng)
    inner = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    for q, col in enumerate((3, 4, 5, 6, 8)):         # rho u v w p
        assert np.array_equal(a[q][inner], vals[..., col])
        ghost = a[q].copy()
        ghost[inner] = np.nan
        assert np.all(np.isnan(ghost))                 # nothing outside the interior is written


def test_read_tecplot_short_file_fails(tool, tmp_path):
    write_file(tmp_path / "in.dat", 2, 2, 2, np.random.default_rng(0))
    with pytest.raises(AssertionError, match="fewer than expected"):
        tool("read", tmp_path / "in.dat", 2, 2, 3, 1, tmp_path / "out.raw")


@pytest.mark.parametrize("src,tgt", [((8, 6, 4), (4, 3, 2)), ((3, 4, 2), (5, 4, 3)), ((7, 5, 1), (3, 2, 1))])
def test_downsample_rule(tool, tmp_path, src, tgt):
    snx, sny, snz = src
    nx, ny, nz = tgt
    ng = 1
    vals = write_file(tmp_path / "in.dat", snx, sny, snz, np.random.default_rng(5))
    tool("downsample", tmp_path / "in.dat", snx, sny, snz, nx, ny, nz, ng, tmp_path / "out.raw")
    a = load_raw(tmp_path / "out.raw", nx, ny, nz, ng)

    def m(g, tn, sn):          # nearest sample, ties away from zero (llround)
        if tn <= 1:
            return 0
        pos = g * sn / tn
        r = int(np.floor(pos + 0.5))
        return min(max(r, 0), sn - 1)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                s = vals[m(k, nz, snz), m(j, ny, sny), m(i, nx, snx)]
                got = a[:, ng + k, ng + j, ng + i]
                assert np.array_equal(got, s[[3, 4, 5, 6, 8]])


def test_spectrum_file(tool, tmp_path):
    ek = [0.0, 0.123456789, 1.0 / 3.0, 2.5e-17, 12345678.9]
    out = tool("spectrum", tmp_path / "s.dat", *["%.17g" % e for e in ek])
    text = (tmp_path / "s.dat").read_text().splitlines()
    assert text == ["0 0", "1 0.123457", "2 0.333333", "3 2.5e-17", "4 1.23457e+07"]
    back = np.array([[float(v) for v in line.split()] for line in out.splitlines()])
    assert np.array_equal(back[:, 0], np.arange(5))
    assert np.array_equal(back[:, 1], [0.0, 0.123457, 0.333333, 2.5e-17, 1.23457e+07])


def test_turbulence_stats_file(tool, tmp_path):
    # E = (0, 0.5, 0.125): sum k^2 E = 1, so eps = 2 nu with nu = mu/rho = 0.01
    tool("spectrum", tmp_path / "s.dat", 0.0, 0.5, 0.125)
    args = dict(ke=0.3, tke=0.29, u_rms=0.4, taylor_li=0.7, re_lambda=28.0, mach_t=0.2, mu=0.02, rho=2.0)
    for t in (0.0, 0.5):
        tool("stats", tmp_path / "s.dat", tmp_path / "ts.dat", t, *args.values())
    lines = (tmp_path / "ts.dat").read_text().splitlines()
    assert lines[0] == ("current_time Kinetic Energy TKE u_rms Dissipation Taylor Kol_scale Kol_vel Kol_time "
                        "Taylor_Li Re_lambda Mach_t")
    assert len(lines) == 3
    nu, eps = 0.01, 0.02
    want = [0.5, 0.3, 0.29, 0.4, eps, np.sqrt(15 * nu * 0.16 / eps), (nu ** 3 / eps) ** 0.25, (nu * eps) ** 0.25,
            (nu / eps) ** 0.5, 0.7, 28.0, 0.2]
    assert lines[2] == " ".join("%.8e" % v for v in want)


def test_residual_file(tool, tmp_path):
    p = tmp_path / "r.dat"
    tool("resid", p, 1, 1, 2, 3, 4, 5, 0.25)
    tool("resid", p, 10, 1e-3, 2e-3, 3e-3, 4e-3, 5e-3, 0.125)
    lines = p.read_text().splitlines()
    assert lines == ['TITLE = "Residuals vs Step"',
                     'VARIABLES = "Step" "Res_rho" "Res_rhou" "Res_rhov" "Res_rhow" "Res_E" "Etot"',
                     "1 1.00000000e+00 2.00000000e+00 3.00000000e+00 4.00000000e+00 5.00000000e+00 2.50000000e-01",
                     "10 1.00000000e-03 2.00000000e-03 3.00000000e-03 4.00000000e-03 5.00000000e-03 1.25000000e-01"]
    tool("resid", p, 1, 0, 0, 0, 0, 0, 0)          # step 1 starts the file again
    assert len(p.read_text().splitlines()) == 3


def test_restart_round_trip(tool, tmp_path):
    nx, ny, nz, ng = 6, 5, 4, 2
    path = tmp_path / "r.bin"
    tool("rwrite", path, nx, ny, nz, ng, 1, 123456789012, 0.1 + 0.2, 0.375)
    raw = path.read_bytes()
    assert raw[:8] == b"NS3DRST2" and len(raw) == 8 + 16 + 8 + 8 + 8 + 5 * nx * ny * nz * 8
    head = tool("rread", path, nx, ny, nz, ng, tmp_path / "out.raw").split()
    assert head[:4] == [str(nx), str(ny), str(nz), "1"] and int(head[4]) == 123456789012
    assert float(head[5]) == 0.1 + 0.2 and float(head[6]) == 0.375
    a = load_raw(tmp_path / "out.raw", nx, ny, nz, ng)
    inner = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for q in range(5):
        assert np.array_equal(a[q][inner], pattern(q, i, j, k) + 1e-13 * q)
        ghost = a[q].copy()
        ghost[inner] = np.nan
        assert np.all(np.isnan(ghost))


def test_restart_reads_the_earlier_layout(tool, tmp_path):
    """A file of the first layout (magic NS3DRST1, no next output time)
    reads with the schedule unknown (NaN)."""
    nx, ny, nz, ng = 3, 2, 2, 1
    path = tmp_path / "r.bin"
    tool("rwrite", path, nx, ny, nz, ng, 0, 7, 0.5, 0.75)
    raw = path.read_bytes()
    old = b"NS3DRST1" + raw[8:40] + raw[48:]
    (tmp_path / "old.bin").write_bytes(old)
    head = tool("rread", tmp_path / "old.bin", nx, ny, nz, ng, tmp_path / "o.raw").split()
    assert int(head[4]) == 7 and float(head[5]) == 0.5 and head[6] == "nan"
    a = load_raw(tmp_path / "o.raw", nx, ny, nz, ng)
    b_path = tmp_path / "o2.raw"
    tool("rread", path, nx, ny, nz, ng, b_path)
    b = load_raw(b_path, nx, ny, nz, ng)
    inner = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    for q in range(5):
        assert np.array_equal(a[q][inner], b[q][inner])


def test_restart_rejects_wrong_size_and_truncation(tool, tmp_path):
    path = tmp_path / "r.bin"
    tool("rwrite", path, 6, 5, 4, 2, 0, 3, 1.5, 2.0)
    with pytest.raises(AssertionError, match="holds a 6x5x4 grid"):
        tool("rread", path, 6, 5, 5, 2, tmp_path / "o.raw")
    path.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(AssertionError, match="truncated"):
        tool("rread", path, 6, 5, 4, 2, tmp_path / "o.raw")
    (tmp_path / "x.bin").write_bytes(b"NOTARST1" + bytes(64))
    with pytest.raises(AssertionError, match="not a restart file"):
        tool("rread", tmp_path / "x.bin", 6, 5, 4, 2, tmp_path / "o.raw")
