"""Build the gfx950 shared library libm2c_hip.so in-tree with hipcc.

The fused stage kernel is compiled once per (flux, equation-of-state) variant from
csrc/stage_inst.hip; the objects are built in parallel and linked with the context code.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libm2c_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-add contraction, so double-precision results equal the
# reference's x86-64 arithmetic operation for operation.
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]

# (tag, flux code, eos kind, arithmetic type) -- flux codes: ROE 0, LLF 1, HLLC 2, GODUNOV 3; eos kind -1 = run-time
# dispatch; float = the mixed-precision form
VARIANTS = [("hllc_sg", 2, 0, "double"), ("hllc_gen", 2, -1, "double"), ("roe_gen", 0, -1, "double"),
            ("llf_gen", 1, -1, "double"), ("hllc_sg_f32", 2, 0, "float"), ("hllc_gen_f32", 2, -1, "float"),
            ("roe_gen_f32", 0, -1, "float"), ("llf_gen_f32", 1, -1, "float"), ("god_gen", 3, -1, "double"),
            ("hllc_sg_mp", 2, 0, "double"), ("hllc_sg_mp_f32", 2, 0, "float")]
# variants that keep the material-interface code with a compile-time equation of state
MP_VARIANTS = {"hllc_sg_mp", "hllc_sg_mp_f32"}
# The mixed-form variants (FP32 reconstruction and fluxes) take the hardware's FP32 reciprocal and
# square root instead of the correctly rounded sequences; FP64 code is not affected by this flag.
# The batched stage units holding FP32 stage bodies take it too (batched and single runs stay equal).
FP32_FAST = ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]
# The FP64 stage kernels are bound by the latency of their dependent FP64 chains at two waves per
# SIMD; the scheduler that maximises instruction-level parallelism shortens the two-material
# launch by ~2 % (rocprofv3, 384x192x192 shock-bubble) and leaves the results unchanged.
ILP_SCHED = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
# The mixed-form two-material unit runs three 256-thread blocks per CU (168 VGPRs, a few hundred
# bytes of spills): 2.02 -> 1.96 ms per stage at the 384x192x192 shock-bubble. The single-gas FP32
# unit is faster at two (2.20 against 2.35 ms at 256^3).
BLOCKS3 = ["-DM2C_STAGE_BLOCKS=3"]
# the batched runner: its host unit, and one unit per stage variant holding both tile shapes
BATCH_VARIANTS = [("hllc_sg", 2, 0, "double", 0), ("hllc_gen", 2, -1, "double", 1), ("roe_gen", 0, -1, "double", 1),
                  ("llf_gen", 1, -1, "double", 1), ("hllc_sg_f32", 2, 0, "float", 0),
                  ("hllc_gen_f32", 2, -1, "float", 1), ("hllc_sg_mp", 2, 0, "double", 1),
                  ("hllc_sg_mp_f32", 2, 0, "float", 1)]
UNITS = [("m2c_hip", "m2c_hip.hip", []), ("multiphase", "multiphase.hip", []), ("iface_multi", "iface_multi.hip", []),
         ("batch", "batch.hip", []),
         ("mlp", "mlp.hip", [])] + [
    (f"bstage_{t}", "batch_stage.hip", [f"-DM2C_FLUX={f}", f"-DM2C_EK={e}", f"-DM2C_TAG={t}", f"-DM2C_TC={tc}",
                                        f"-DM2C_MP={mp}"] + (FP32_FAST if tc == "float" else []))
    for t, f, e, tc, mp in BATCH_VARIANTS] + [
    (f"stage_{t}", "stage_inst.hip", [f"-DM2C_FLUX={f}", f"-DM2C_EK={e}", f"-DM2C_TAG={t}", f"-DM2C_TC={tc}"]
     + (["-DM2C_MP=1"] if t in MP_VARIANTS else []) + (FP32_FAST if tc == "float" else []) + (ILP_SCHED if t in ("hllc_sg", "hllc_sg_mp") else [])
     + (BLOCKS3 if t == "hllc_sg_mp_f32" else []))
    for t, f, e, tc in VARIANTS]
# host-only C++ (no device code), compiled with the host compiler
HOST_UNITS = [("input_deck", "input_deck.cpp"), ("output", "output.cpp"), ("mesh", "mesh.cpp"),
              ("embedded_walls", "embedded_walls.cpp")]
CXX = os.environ.get("CXX", "g++")


def _compile(unit, verbose):
    name, src, defs = unit
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + FLAGS + defs + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src} ({name})")
    return obj


def _compile_host(unit, verbose):
    name, src = unit
    obj = os.path.join(OBJ, name + ".o")
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{CXX} failed on {src}")
    return obj


def build(verbose=False, jobs=None, only=None):
    """Compile every unit (or, with `only`, just the named units, reusing the other objects of an
    earlier build) and link libm2c_hip.so."""
    os.makedirs(OBJ, exist_ok=True)
    todo = [u for u in UNITS if only is None or u[0] in only]
    jobs = jobs or max(1, min(len(todo), os.cpu_count() or 2))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda u: _compile(u, verbose), todo))
    for u in HOST_UNITS:
        if only is None or u[0] in only:
            _compile_host(u, verbose)
    objs = [os.path.join(OBJ, u[0] + ".o") for u in UNITS + HOST_UNITS]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError("objects missing (run a full build first): " + ", ".join(missing))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed for libm2c_hip.so")
    os.replace(tmp, OUT)
    build_host_demo(verbose)
    return OUT


DEMO = os.path.join(HERE, "m2c_demo")


def build_host_demo(verbose=False):
    """The C++ host-layer driver (host/m2c_demo.cpp), linked against libm2c_hip.so."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "host", "m2c_demo.cpp"),
           "-L" + HERE, "-lm2c_hip", "-Wl,-rpath,$ORIGIN", "-o", DEMO]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed on host/m2c_demo.cpp")
    return DEMO


if __name__ == "__main__":
    only = None
    if "--only" in sys.argv:
        only = set(sys.argv[sys.argv.index("--only") + 1].split(","))
    print(build(verbose="-v" in sys.argv, only=only))
