"""The C-ABI library loads and exports every entry point its header declares
(no GPU calls), host-side argument checks, and the `pp` command line."""
import ctypes
import os
import re
import subprocess

import pytest

import pphpc_amd as pp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pphpc.h")
PP_EXE = os.path.join(ROOT, "pphpc_amd", "bin", "pp")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pp_\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("pp_load_params", "pp_stats_save", "pp_sim_create",
                 "pp_sim_init", "pp_sim_step", "pp_sim_run", "pp_sim_destroy",
                 "pp_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pp.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", pp.library_path()],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (pp_\w+)", out))
    assert set(declared_functions()) <= exported
    # C linkage only: no mangled C++ names among the pp_ entry points
    assert not re.search(r" T _Z\w*pp_sim", out)


def test_library_is_gfx950_code_object(tmp_path):
    """The embedded offload bundle of the library targets gfx950 only."""
    tools = "/opt/rocm/llvm/bin"
    fat = str(tmp_path / "fatbin.bin")
    r = subprocess.run([f"{tools}/llvm-objcopy",
                        f"--dump-section=.hip_fatbin={fat}",
                        pp.library_path(), str(tmp_path / "copy.so")],
                       capture_output=True)
    if r.returncode != 0:
        pytest.skip("llvm-objcopy not available")
    out = subprocess.run([f"{tools}/clang-offload-bundler", "--list",
                          "--type=o", f"--input={fat}"], capture_output=True,
                         text=True, check=True).stdout.split()
    devices = [t for t in out if t.startswith("hip")]
    assert devices == ["hipv4-amdgcn-amd-amdhsa--gfx950"]


@pytest.mark.parametrize("field,value", [
    ("grid_x", 0), ("grid_y", 0), ("grass_restart", 0),
])
def test_create_rejects_invalid_parameters(field, value):
    """Checked on the host before any HIP call."""
    d = dict(init_sheep=10, sheep_gain_from_food=4,
             sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
             init_wolves=5, wolves_gain_from_food=20,
             wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
             grass_restart=10, grid_x=10, grid_y=10, iters=10)
    d[field] = value
    with pytest.raises(pp.PPError) as e:
        pp.Simulation(pp.Params.from_values(**d), capacity=100)
    assert e.value.code == pp.PP_INVALID_ARGS


def test_create_rejects_capacity_below_population():
    d = dict(init_sheep=100, sheep_gain_from_food=4,
             sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
             init_wolves=50, wolves_gain_from_food=20,
             wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
             grass_restart=10, grid_x=10, grid_y=10, iters=10)
    with pytest.raises(pp.PPError) as e:
        pp.Simulation(pp.Params.from_values(**d), capacity=149)
    assert e.value.code == pp.PP_INVALID_ARGS
    with pytest.raises(pp.PPError) as e:
        pp.Simulation(pp.Params.from_values(**d), capacity=1 << 29)
    assert e.value.code == pp.PP_INVALID_ARGS


def _pp(*args, cwd=None):
    return subprocess.run([PP_EXE, *args], capture_output=True, text=True,
                          cwd=cwd, timeout=60)


def test_cli_help():
    r = _pp("-h")
    assert r.returncode == 0 and "--params" in r.stdout


@pytest.mark.parametrize("args,code", [
    (["--bogus"], pp.PP_UNKNOWN_ARGS),
    (["extra-positional"], pp.PP_UNKNOWN_ARGS),
    (["-r", "12x"], pp.PP_INVALID_ARGS),
    (["-p", "/nonexistent.txt"], pp.PP_UNABLE_TO_OPEN_PARAMS_FILE),
])
def test_cli_error_exit_codes(args, code):
    """Exit status = the PP_* code, as the reference's main returns it."""
    r = _pp(*args)
    assert r.returncode == code & 0xFF, r.stderr


def test_cli_invalid_params_file(tmp_path):
    bad = tmp_path / "config.txt"
    bad.write_text("INIT_SHEEP = 3\nINIT_SHEEP = 4\n")
    r = _pp(cwd=str(tmp_path))          # default -p config.txt
    assert r.returncode == pp.PP_INVALID_PARAMS_FILE & 0xFF


@pytest.mark.parametrize("args", [
    (0, 100, 4, 10), (1 << 29, 100, 4, 10), (100, 0, 4, 10),
    (100, 1 << 32, 4, 10), (100, 100, 0, 10), (100, 100, 4, 0),
])
def test_batch_create_rejects_invalid_limits(args):
    """Checked on the host before any HIP call."""
    with pytest.raises(pp.PPError) as e:
        pp.Batch(*args)
    assert e.value.code == pp.PP_INVALID_ARGS
    assert pp.Batch.footprint(*args) == 0


def test_batch_footprint_counts_every_buffer():
    """33 bytes per agent (SoA 9 + three pair buffers 24), the grass (1 byte
    per cell for a restart <= 191), rings, histograms and stats rings."""
    A, C = 1_000_000, 4_000_000
    f = pp.Batch.footprint(A, C, 8, 10)
    assert 33 * A + C < f < 33 * A + C + 8 * 1024 * 56 + (40 << 20)
    assert pp.Batch.footprint(A, C, 8, 300) - f == C + 8 * 290 * 8


def test_step_kernels_do_not_spill(tmp_path):
    """The step kernels keep everything in registers: a spill to scratch
    memory slows a kernel many times over while every result stays the
    same, so no parity test sees it. Compiles the library's HIP source with
    the compiler's resource remarks and checks each step kernel's scratch
    (the per-bucket sort keeps the 36 bytes a lane it has always had)."""
    import re
    src = os.path.join(ROOT, "pphpc_amd", "csrc", "pp_sim.hip")
    r = subprocess.run(
        ["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "-fPIC",
         "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-c",
         "-o", str(tmp_path / "x.o"), src,
         "-Rpass-analysis=kernel-resource-usage"],
        capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    scratch, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            scratch[name] = int(m.group(1))
    limits = {"k_act": 0, "k_act_mixed": 0, "k_emit_move": 0,
              "k_act_scan": 0, "k_sort_down": 0,
              "k_sort_up": 0, "k_bucket_sort": 36}
    seen = set()
    for mangled, size in scratch.items():
        for k, lim in limits.items():
            if re.search(r"\d%s[IE]" % re.escape(k), mangled):
                seen.add(k)
                assert size <= lim, f"{k} ({mangled}) spills {size} B/lane"
    assert seen == set(limits), set(limits) - seen
