"""Run the library's host code on the CPU against a stand-in HIP runtime
(fakehip.c): device buffers live in a numpy arena and every kernel launch is
executed by interpreting the kernel's compiled gfx950 assembly (gcnemu.py),
with every memory access checked against the live allocations.

Test infrastructure only (tests/test_isa_compiled_kernels.py). Create a Sim()
before the first use of cgfdm_amd.solver in the process: it loads the stand-in
runtime under the soname the library links against.
"""
import ctypes
import glob
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gcnemu as G  # noqa: E402

REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "cgfdm3d-eqr-sc_amd", "csrc")
BUILD = os.path.join(HERE, "_build")
# extra preprocessor definitions of the kernel build under test (space separated,
# e.g. CGFDM_ISA_DEFINES="CGFDM_NT_STREAMS"), compiled into their own folder
DEFINES = os.environ.get("CGFDM_ISA_DEFINES", "").split()
ASMDIR = os.path.join(BUILD, "asm" + "".join("_" + d for d in DEFINES))
FAKEHIP = os.path.join(BUILD, "libamdhip64.so.7")


def build_fakehip():
    """The stand-in runtime, exporting the versioned HIP symbols the library imports."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "fakehip.c")
    if os.path.exists(FAKEHIP) and os.path.getmtime(FAKEHIP) > os.path.getmtime(src):
        return FAKEHIP
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-Wl,--version-script=" + os.path.join(HERE, "hip.map"),
                           "-Wl,-soname,libamdhip64.so.7", "-o", FAKEHIP, src])
    return FAKEHIP
SOURCES = ["setup.hip", "step_fp32.hip", "step_fp64.hip", "step_fp16.hip", "step_bf16.hip", "aweno_fp32.hip",
           "aweno_fp64.hip", "adjoint_fp32.hip", "adjoint_fp64.hip"]


def build_asm(force=False):
    os.makedirs(ASMDIR, exist_ok=True)
    procs = []
    for s in SOURCES:
        out = os.path.join(ASMDIR, s.replace(".hip", ".s"))
        src = os.path.join(CSRC, s)
        deps = [src] + glob.glob(os.path.join(CSRC, "*.h"))
        if not force and os.path.exists(out) and os.path.getmtime(out) > max(os.path.getmtime(d) for d in deps):
            continue
        cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature",
               "-Xclang", "-packed-fp32-ops", "-I", CSRC, "-I", os.path.join(REPO, "include"), "--cuda-device-only",
               "-S", src, "-o", out] + ["-D" + d for d in DEFINES]
        procs.append(subprocess.Popen(cmd, stderr=subprocess.DEVNULL))
    for p in procs:
        assert p.wait() == 0


class Sim:
    def __init__(self, arena_bytes=1 << 30, verbose=False):
        build_asm()
        self.texts = [open(f).read() for f in sorted(glob.glob(os.path.join(ASMDIR, "*.s")))]
        self.kernels = {}
        self.mem = np.zeros(arena_bytes, np.uint8)
        self.ar = G.Arena.__new__(G.Arena)
        self.ar.mem = self.mem
        self.ar.valid = np.zeros(arena_bytes // 4, bool)
        self.ar.BASE = self.mem.ctypes.data
        self.ar.GAP = 1 << 16
        self.ar.top = self.ar.GAP
        self.ar.bufs = []
        self.verbose = verbose
        self.nlaunch = 0
        self.twall = 0.0
        self.fault = None
        self.lib = ctypes.CDLL(build_fakehip(), mode=ctypes.RTLD_GLOBAL)
        A = ctypes.CFUNCTYPE(ctypes.c_uint64, ctypes.c_uint64)
        F = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint64)
        L = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_char_p, *([ctypes.c_uint32] * 6), ctypes.POINTER(ctypes.c_void_p),
                             ctypes.c_uint64)
        self._cbs = (A(self._alloc), F(self._free), L(self._launch))
        self.lib.fakehip_set_callbacks(*self._cbs)

    # ---------------------------------------------------------------- memory
    def _alloc(self, n):
        ar = self.ar
        n4 = max(4, (int(n) + 3) // 4 * 4)
        off = (ar.top + 255) // 256 * 256
        if off + n4 + ar.GAP > ar.mem.size:
            return 0
        ar.top = off + n4 + ar.GAP
        ar.valid[off // 4:(off + n4) // 4] = True
        ar.bufs.append((ar.BASE + off, n4, f"buf{len(ar.bufs)}({n})"))
        # fresh device memory holds garbage: NaN pattern
        ar.mem[off:off + n4] = 0xFF
        return ar.BASE + off

    def _free(self, addr):
        ar = self.ar
        for q, (b, n, nm) in enumerate(ar.bufs):
            if b == addr:
                o = b - ar.BASE
                ar.valid[o // 4:(o + n) // 4] = False
                ar.bufs[q] = (b, n, nm + "[freed]")
                return 0
        print("fakehip: free of unknown pointer", hex(addr), flush=True)
        return 1

    # ---------------------------------------------------------------- kernels
    def kernel(self, name):
        k = self.kernels.get(name)
        if k is None:
            for t in self.texts:
                if "\n" + name + ":" in t:
                    k = G.Kernel(t, name)
                    break
            assert k is not None, "no assembly for " + name
            self.kernels[name] = k
        return k

    def _launch(self, name, gx, gy, gz, bx, by, bz, args, shmem):
        name = name.decode()
        t0 = time.time()
        try:
            k = self.kernel(name)
            assert by == 1 and bz == 1, "1-D blocks only"
            ka = bytearray(max(256, k.desc[".amdhsa_kernarg_size"]))
            i = 0
            for off, size, kind in k.args:
                if kind.startswith("hidden_"):
                    val = {"hidden_block_count_x": gx, "hidden_block_count_y": gy, "hidden_block_count_z": gz,
                           "hidden_group_size_x": bx, "hidden_group_size_y": by, "hidden_group_size_z": bz,
                           "hidden_grid_dims": 1 + (gy > 1 or gz > 1) + (gz > 1)}.get(kind, 0)
                    ka[off:off + size] = int(val).to_bytes(size, "little")
                else:
                    ka[off:off + size] = ctypes.string_at(args[i], size)
                    i += 1
            kaddr = self._alloc(len(ka))
            self.ar.mem[kaddr - self.ar.BASE:kaddr - self.ar.BASE + len(ka)] = np.frombuffer(bytes(ka), np.uint8)
            lds = k.desc.get(".amdhsa_group_segment_fixed_size", 0) > 0
            for z in range(gz):
                for y in range(gy):
                    for x in range(gx):
                        if lds:
                            k.run_block(self.ar, kaddr, (x, y, z), bx)
                        else:
                            for w0 in range(0, bx, 64):
                                k.run_wave(self.ar, kaddr, (x, y, z), w0, min(64, bx - w0))
            self._free(kaddr)
        except G.OOB as e:
            self.fault = f"{name} grid ({gx},{gy},{gz}) block {bx}: {e}"
            print("FAULT", self.fault, flush=True)
            return 1
        except Exception as e:  # interpreter gap, not a kernel fault
            import traceback
            traceback.print_exc()
            self.fault = f"INTERPRETER {type(e).__name__}: {e} in {name}"
            print("FAULT", self.fault, flush=True)
            return 1
        self.nlaunch += 1
        self.twall += time.time() - t0
        if self.verbose:
            print(f"  launch {self.nlaunch}: {name[:60]} ({gx},{gy},{gz})x{bx} {time.time() - t0:.1f}s", flush=True)
        return 0
