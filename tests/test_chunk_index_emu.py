"""Host-side emulation of the production substep's index arithmetic, under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/emu/chunk_index_emu.cpp).

The emulation runs the kernels' own index functions (mpm_index.h) and tables
(mpm_tables.cpp) over the oracle's trajectory, which the library reproduces
bit for bit, and checks every index against the buffers Batch::init
allocates, that every launch shape's chunk walk visits each tile chunk once,
that the scan's chunks tile the slots and hold only their tile's particles,
and that the permutation is a bijection.  The oracle itself is built with the
same sanitizers.  CPU only: nothing here touches a GPU."""
import os
import subprocess

import numpy as np
import pytest

from mlsmpm import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = tmp_path_factory.mktemp("emu")
    obj = str(out / "mpm_oracle.o")
    exe = str(out / "chunk_index_emu")
    subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    *SAN, "-c", os.path.join(ROOT, "oracle", "mpm_oracle.c"), "-o", obj],
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fopenmp", *SAN, "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "webgpu-ocean_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "emu", "chunk_index_emu.cpp"),
                    os.path.join(ROOT, "webgpu-ocean_amd", "csrc", "mpm_tables.cpp"), obj,
                    "-o", exe, "-lm"], check=True)
    return out, exe


def run(emu, rec, box, substeps, ppt, resize=()):
    out, exe = emu
    path = str(out / f"particles_{rec.shape[0]}_{ppt}.bin")
    np.ascontiguousarray(rec, dtype=np.float32).tofile(path)
    args = [exe, path, str(rec.shape[0]), *[str(float(b)) for b in box], str(substeps), str(ppt),
            "8", *[f"{s}:{x}" for s, x in resize]]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="8")
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:] + r.stdout[-1000:]
    stats = dict(kv.split("=") for kv in r.stdout.split()[1:])
    return {k: int(v) for k, v in stats.items()}


def test_dam_break_trajectory_with_box_changes(emu):
    # 40k dam break, 240 substeps; the box shrinks to 48 in x at substep 120
    rec, box = scenes.config0()
    st = run(emu, rec, box, 240, 2, resize=[(120, 48.0)])
    assert st["substeps"] == 240 and st["checks"] > 10**7
    assert st["max_chunks"] <= st["cap_chunks"]


@pytest.mark.parametrize("ppt", [1, 4])
def test_far_moves_and_wild_particles(emu, ppt):
    # Fast particles jump tiles in one substep (final ranks through the key
    # counters); a real box larger than the init box lets g2p clamp particles
    # outside the tame range (the wild key and its chunks), then it shrinks.
    box = [40.0, 36.0, 44.0]
    rec = scenes.small_scene(20000, box, seed=21, velocity=1.5, affine=0.4)
    rec[:200, 4:7] = np.random.default_rng(3).choice([-400.0, 400.0], size=(200, 3))
    st = run(emu, rec, box, 60, ppt, resize=[(10, 47.0), (30, 33.5)])
    assert st["final_ranks"] > 200      # far moves and wild particles took final ranks
    assert st["wild_substeps"] >= 5     # substeps whose g2p sent particles to the wild key
