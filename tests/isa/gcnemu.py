"""Interpreter of the gfx950 (CDNA) instruction subset the library's kernels
compile to: scalar and vector ALU (FP32/FP64/integer, 16/24-bit and SDWA
forms), exec masking, SGPR spills to VGPR lanes, AGPR moves, global loads,
stores and float atomics, LDS and barriers. It executes a kernel's assembly
text wave by wave over a flat memory arena and raises OOB on any access
outside the live allocations.

Test infrastructure only. Semantics that do not affect addressing are
approximated where noted (division scaling steps return the unscaled value;
the carry-out of 64-bit multiply-adds is reported as zero); results are
checked against the CPU oracle, so an approximation that mattered would show.
"""
import re
import struct
import numpy as np

NL = 64
M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1
LANES = np.arange(NL, dtype=np.uint64)
VCC, EXEC, M0 = 106, 126, 124

INLINE_F = {"0.5": 0.5, "-0.5": -0.5, "1.0": 1.0, "-1.0": -1.0, "2.0": 2.0, "-2.0": -2.0,
            "4.0": 4.0, "-4.0": -4.0, "0.15915494": 0.15915494309189532}


class OOB(Exception):
    pass


class Arena:
    """All device buffers in one byte array with unmapped gaps between them."""

    BASE = 0x7F0000000000
    GAP = 1 << 16

    def __init__(self, size=1 << 28):
        self.mem = np.zeros(size, np.uint8)
        self.valid = np.zeros(size // 4, bool)
        self.top = self.GAP
        self.bufs = []

    def alloc(self, nbytes, name):
        nbytes = max(int(nbytes), 4)
        off = (self.top + 255) // 256 * 256
        self.top = off + nbytes + self.GAP
        assert self.top < self.mem.size, "arena too small"
        assert nbytes % 4 == 0
        self.valid[off // 4:(off + nbytes) // 4] = True
        self.bufs.append((self.BASE + off, nbytes, name))
        return self.BASE + off

    def name_of(self, addr):
        for b, n, nm in self.bufs:
            if b - self.GAP // 2 <= addr < b + n + self.GAP // 2:
                return f"{nm}[{addr - b:+d}] (size {n})"
        return "unmapped"

    def write(self, addr, data: bytes):
        o = addr - self.BASE
        self.mem[o:o + len(data)] = np.frombuffer(data, np.uint8)

    def read(self, addr, n):
        o = addr - self.BASE
        return self.mem[o:o + n].tobytes()

    def check(self, addrs, nbytes, what):
        addrs = np.asarray(addrs, np.uint64)
        if addrs.size == 0:
            return
        off = addrs.astype(np.int64) - self.BASE
        bad = (off < 0) | (off + nbytes > self.mem.size) | (off % min(nbytes, 4) != 0)
        if not bad.any():
            for k in range(0, max(nbytes, 4) if nbytes >= 4 else 1, 4):
                bad |= ~self.valid[np.clip((off + k) // 4, 0, self.valid.size - 1)]
        if bad.any():
            a = int(addrs[np.argmax(bad)])
            raise OOB(f"{what}: {int(bad.sum())} lanes out of bounds, first 0x{a:x} -> {self.name_of(a)}")

    def gather32(self, addrs):
        off = (addrs.astype(np.int64) - self.BASE) // 4
        return self.mem.view(np.uint32)[off]

    def gather16(self, addrs):
        off = (addrs.astype(np.int64) - self.BASE) // 2
        return self.mem.view(np.uint16)[off]

    def scatter16(self, addrs, vals):
        off = (addrs.astype(np.int64) - self.BASE) // 2
        self.mem.view(np.uint16)[off] = vals

    def scatter32(self, addrs, vals):
        off = (addrs.astype(np.int64) - self.BASE) // 4
        self.mem.view(np.uint32)[off] = vals


def parse_kernel(text, name):
    """Instructions of one kernel: list of (op, [operand strings], mods) and label map."""
    lines = text.split("\n")
    start = None
    for n, l in enumerate(lines):
        if l.startswith(name + ":"):
            start = n
            break
    assert start is not None, name
    ins, labels = [], {}
    for l in lines[start + 1:]:
        s = l.split(";")[0].strip()
        if not s:
            continue
        if s.startswith(".LBB") and s.endswith(":"):
            labels[s[:-1]] = len(ins)
            continue
        if s.startswith("."):
            if s.startswith(".Lfunc_end"):
                break
            continue
        parts = s.split(None, 1)
        op = parts[0]
        rest = parts[1] if len(parts) > 1 else ""
        mods = {}
        for key in ("dst_sel", "dst_unused", "src0_sel", "src1_sel"):
            m = re.search(r"\s" + key + r":(\w+)", " " + rest)
            if m:
                mods[key] = m.group(1)
                rest = re.sub(r"\s*" + key + r":\w+", "", rest)
        for key in ("op_sel_hi", "op_sel"):
            m = re.search(r"\s" + key + r":\[([0-9,]+)\]", " " + rest)
            if m:
                mods[key] = [int(x) for x in m.group(1).split(",")]
                rest = re.sub(r"\s*" + key + r":\[[0-9,]+\]", "", rest)
        for key in ("offset0", "offset1"):
            m = re.search(r"\s" + key + r":(\w+)", " " + rest)
            if m:
                mods[key] = int(m.group(1), 0)
                rest = re.sub(r"\s*" + key + r":\w+", "", rest)
        m = re.search(r"\soffset:(-?\w+)", " " + rest)
        if m:
            mods["offset"] = int(m.group(1), 0)
            rest = re.sub(r"\s*offset:-?\w+", "", rest)
        for flag in ("glc", "slc", "nt", "sc0", "sc1"):
            if re.search(r"\b" + flag + r"\b", rest):
                mods[flag] = True
                rest = re.sub(r"\s*\b" + flag + r"\b", "", rest)
        ops = [o.strip() for o in rest.split(",")] if rest else []
        ins.append((op, ops, mods))
        if op == "s_endpgm" and False:
            break
    return ins, labels


def kernel_descriptor(text, name):
    m = re.search(r"\.amdhsa_kernel " + re.escape(name) + r"\n(.*?)\.end_amdhsa_kernel", text, re.S)
    d = {}
    for l in m.group(1).split("\n"):
        p = l.split()
        if len(p) == 2:
            d[p[0]] = int(p[1])
    return d


def kernel_args(text, name):
    """(offset, size, value_kind) of the kernel's arguments from the metadata."""
    i = text.index(".name:           " + name)
    j = text.rfind(".args:", 0, i)
    block = text[j:i]
    out = []
    for m in re.finditer(r"\.offset:\s+(\d+)\s+\.size:\s+(\d+)\s+\.value_kind:\s+(\w+)", block):
        out.append((int(m.group(1)), int(m.group(2)), m.group(3)))
    return out


import functools


@functools.lru_cache(maxsize=None)
def vreg(t):
    """('v'|'a', first index) of a vector register operand, or None."""
    m = re.fullmatch(r"([va])\[(\d+):(\d+)\]", t) or re.fullmatch(r"([va])(\d+)()", t)
    return (m.group(1), int(m.group(2))) if m else None


@functools.lru_cache(maxsize=None)
def sreg(t):
    if t in ("vcc", "vcc_lo"):
        return VCC, 1 if t == "vcc_lo" else 2
    if t == "vcc_hi":
        return VCC + 1, 1
    if t in ("exec", "exec_lo"):
        return EXEC, 1 if t == "exec_lo" else 2
    if t == "exec_hi":
        return EXEC + 1, 1
    if t == "m0":
        return M0, 1
    m = re.fullmatch(r"s\[(\d+):(\d+)\]", t)
    if m:
        a, b = int(m.group(1)), int(m.group(2))
        return a, b - a + 1
    m = re.fullmatch(r"s(\d+)", t)
    if m:
        return int(m.group(1)), 1
    return None


@functools.lru_cache(maxsize=None)
def negsplit(t):
    if t.startswith("-") and not re.fullmatch(r"-\d+", t) and t not in INLINE_F:
        return True, t[1:]
    return False, t


class Wave:
    def __init__(self, arena):
        self.ar = arena
        self.S = np.zeros(128, np.uint64)  # 32-bit values kept in uint64
        self.V = np.zeros((256, NL), np.uint32)
        self.A = np.zeros((256, NL), np.uint32)
        self.scc = 0

    # ---- scalar helpers -------------------------------------------------
    def sidx(self, t):
        return sreg(t)

    def sget(self, t, width, fp=False):
        """Scalar operand as a Python int of `width` dwords (fp: f64 literal rule)."""
        neg, t = negsplit(t)
        r = self.sidx(t)
        if r is not None:
            i, n = r
            v = 0
            for k in range(width):
                v |= (int(self.S[i + k]) & M32) << (32 * k)
        elif t == "scc":
            v = self.scc
        elif t in INLINE_F:
            f = INLINE_F[t]
            v = struct.unpack("<Q", struct.pack("<d", f))[0] if width == 2 else struct.unpack("<I", struct.pack("<f", f))[0]
            if not fp and width == 1:
                pass
        elif re.fullmatch(r"-?\d+", t):
            v = int(t) & (M64 if width == 2 else M32)
        elif t.startswith("0x"):
            v = int(t, 16)
            if width == 2:
                if fp:
                    v = v << 32
                else:
                    raise NotImplementedError("64-bit integer literal " + t)
        else:
            raise NotImplementedError("scalar operand " + t)
        if neg:
            v ^= 1 << (32 * width - 1)
        return v

    def sset(self, t, width, v):
        i, n = self.sidx(t)
        for k in range(width):
            self.S[i + k] = (v >> (32 * k)) & M32

    def exec_mask(self):
        e = int(self.S[EXEC]) | (int(self.S[EXEC + 1]) << 32)
        return ((np.uint64(e) >> LANES) & np.uint64(1)).astype(bool), e

    def mask_of(self, v):
        return ((np.uint64(v) >> LANES) & np.uint64(1)).astype(bool)

    def bits_of(self, b):
        return int((b.astype(np.uint64) << LANES).sum()) if False else int(np.packbits(b[::-1].astype(np.uint8)).view(">u8")[0])

    # ---- vector helpers -------------------------------------------------
    def vget(self, t, width, fp=False):
        """Per-lane operand as uint64 array (width 1 or 2 dwords)."""
        neg, t = negsplit(t)
        if t.startswith("|") and t.endswith("|"):
            v = self.vget(t[1:-1], width, fp) & np.uint64((1 << (32 * width - 1)) - 1)
            return v ^ np.uint64(1 << (32 * width - 1)) if neg else v
        m = vreg(t)
        if m:
            R = self.V if m[0] == "v" else self.A
            i = m[1]
            v = R[i].astype(np.uint64)
            if width == 2:
                v = v | (R[i + 1].astype(np.uint64) << np.uint64(32))
        else:
            v = np.full(NL, self.sget(t, width, fp), np.uint64)
        if neg:
            v = v ^ np.uint64(1 << (32 * width - 1))
        return v

    def vset(self, t, width, val, mask):
        m = vreg(t)
        R = self.V if m[0] == "v" else self.A
        i = m[1]
        val = np.asarray(val, np.uint64)
        R[i][mask] = (val & np.uint64(M32)).astype(np.uint32)[mask]
        if width == 2:
            R[i + 1][mask] = (val >> np.uint64(32)).astype(np.uint32)[mask]


def f64(u):
    return np.asarray(u, np.uint64).view(np.float64)


def u64(f):
    return np.asarray(f, np.float64).view(np.uint64)


def f32(u):
    return (np.asarray(u, np.uint64) & np.uint64(M32)).astype(np.uint32).view(np.float32)


def u32f(f):
    return np.asarray(f, np.float32).view(np.uint32).astype(np.uint64)


def s32(u):
    return (np.asarray(u, np.uint64) & np.uint64(M32)).astype(np.uint32).view(np.int32).astype(np.int64)


def sx16(x):
    x &= 0xFFFF
    return x - (1 << 16) if x & 0x8000 else x


def sx32(x):
    x &= M32
    return x - (1 << 32) if x & 0x80000000 else x


class Kernel:
    def __init__(self, text, name):
        self.name = name
        self.ins, self.labels = parse_kernel(text, name)
        self.desc = kernel_descriptor(text, name)
        self.args = kernel_args(text, name)
        self.stats = {}

    def init_wave(self, arena, kernarg_addr, wg, tid0, nthreads_valid):
        w = Wave(arena)
        d = self.desc
        assert d.get(".amdhsa_user_sgpr_kernarg_segment_ptr", 0) == 1
        assert d.get(".amdhsa_user_sgpr_dispatch_ptr", 0) == 0 and d.get(".amdhsa_user_sgpr_queue_ptr", 0) == 0
        w.S[0] = kernarg_addr & M32
        w.S[1] = kernarg_addr >> 32
        s = d[".amdhsa_user_sgpr_count"]
        for ax, key in enumerate(("x", "y", "z")):
            if d.get(".amdhsa_system_sgpr_workgroup_id_" + key, 0):
                w.S[s] = wg[ax]
                s += 1
        if d.get(".amdhsa_enable_private_segment", 0):
            # private segment wave offset follows the workgroup ids; scratch is
            # addressed per lane through the wave's own scratch window
            w.S[s] = 0
            s += 1
        assert d.get(".amdhsa_system_vgpr_workitem_id", 0) == 0  # 1-D blocks: v0 = x id
        w.V[0] = (tid0 + np.arange(NL)).astype(np.uint32)
        active = np.arange(NL) < nthreads_valid
        e = int(np.packbits(active[::-1].astype(np.uint8)).view(">u8")[0])
        w.S[EXEC] = e & M32
        w.S[EXEC + 1] = e >> 32
        w.pc = 0
        w.done = False
        w.steps = 0
        w.scratch = {}
        return w

    def run_block(self, arena, kernarg_addr, wg, nthreads, max_steps=4_000_000):
        nwaves = (nthreads + 63) // 64
        lds_size = self.desc.get(".amdhsa_group_segment_fixed_size", 0)
        lds = np.zeros(max(lds_size, 4), np.uint8)
        waves = []
        for q in range(nwaves):
            w = self.init_wave(arena, kernarg_addr, wg, 64 * q, min(64, nthreads - 64 * q))
            w.lds = lds
            w.lds_size = lds_size
            waves.append(w)
        while True:
            for w in waves:
                if w.done:
                    continue
                while True:
                    op, o, mods = self.ins[w.pc]
                    w.steps += 1
                    if w.steps > max_steps:
                        raise RuntimeError("step limit")
                    if op == "s_barrier":
                        w.pc += 1
                        break
                    npc = self.step(w, op, o, mods, w.pc)
                    if npc is None:
                        w.pc += 1
                    elif npc == -1:
                        w.done = True
                        break
                    else:
                        w.pc = npc
            if all(w.done for w in waves):
                return

    def run_wave(self, arena, kernarg_addr, wg, tid0, nthreads_valid, max_steps=2_000_000):
        assert self.desc.get(".amdhsa_group_segment_fixed_size", 0) == 0
        w = self.init_wave(arena, kernarg_addr, wg, tid0, nthreads_valid)
        w.lds = None
        while True:
            op, o, mods = self.ins[w.pc]
            w.steps += 1
            if w.steps > max_steps:
                raise RuntimeError("step limit")
            assert op != "s_barrier"
            npc = self.step(w, op, o, mods, w.pc)
            if npc is None:
                w.pc += 1
            elif npc == -1:
                return w.steps
            else:
                w.pc = npc

    # ------------------------------------------------------------------
    def step(self, w, op, o, mods, pc):
        ar = w.ar
        self.stats[op] = self.stats.get(op, 0) + 1
        # ---------------- program flow
        if op == "s_endpgm":
            return -1
        if op in ("s_waitcnt", "s_nop"):
            return None
        if op == "s_branch":
            return self.labels[o[0]]
        if op.startswith("s_cbranch_"):
            c = op[len("s_cbranch_"):]
            vcc = int(w.S[VCC]) | (int(w.S[VCC + 1]) << 32)
            ex = int(w.S[EXEC]) | (int(w.S[EXEC + 1]) << 32)
            take = {"vccnz": vcc != 0, "vccz": vcc == 0, "scc0": w.scc == 0, "scc1": w.scc == 1,
                    "execz": ex == 0, "execnz": ex != 0}[c]
            return self.labels[o[0]] if take else None
        # ---------------- scalar memory
        if op.startswith("s_load_dword"):
            n = {"s_load_dword": 1, "s_load_dwordx2": 2, "s_load_dwordx4": 4, "s_load_dwordx8": 8,
                 "s_load_dwordx16": 16}[op]
            base = w.sget(o[1], 2)
            off = int(o[2], 0) if len(o) > 2 else 0
            addr = base + off
            ar.check(np.array([addr], np.uint64), 4 * n, f"{op} pc={pc}")
            vals = np.frombuffer(ar.read(addr, 4 * n), np.uint32)
            i, _ = w.sidx(o[0])
            for k in range(n):
                w.S[i + k] = int(vals[k])
            return None
        # ---------------- scalar ALU
        if op.startswith("s_"):
            return self.salu(w, op, o)
        # ---------------- vector memory
        if op.startswith("global_atomic_"):
            return self.vatomic(w, op, o, mods, pc)
        if op.startswith("global_"):
            return self.vmem(w, op, o, mods, pc)
        if op.startswith("ds_"):
            return self.lds(w, op, o, mods, pc)
        if op.endswith("_sdwa"):
            return self.sdwa(w, op[:-5], o, mods)
        if op == "v_fma_mix_f32":
            act, _ = w.exec_mask()
            hi = mods.get("op_sel_hi", [0, 0, 0])
            sel = mods.get("op_sel", [0, 0, 0])
            xs = []
            for k in range(3):
                t = o[k + 1]
                neg, t = negsplit(t)
                absf = t.startswith("|")
                u = w.vget(t.strip("|"), 1)
                if hi[k]:
                    h = ((u >> np.uint64(16 * sel[k])) & np.uint64(0xFFFF)).astype(np.uint16).view(np.float16)
                    x = h.astype(np.float64)
                else:
                    x = f32(u).astype(np.float64)
                if absf:
                    x = np.abs(x)
                if neg:
                    x = -x
                xs.append(x)
            w.vset(o[0], 1, u32f((xs[0] * xs[1] + xs[2]).astype(np.float32)), act)
            return None
        return self.valu(w, op, o)

    def salu(self, w, op, o):
        g = w.sget
        if op == "s_mov_b32":
            w.sset(o[0], 1, g(o[1], 1))
        elif op == "s_mov_b64":
            w.sset(o[0], 2, g(o[1], 2))
        elif op in ("s_and_b64", "s_or_b64", "s_andn2_b64", "s_orn2_b64", "s_and_b32", "s_or_b32"):
            wd = 2 if op.endswith("64") else 1
            a, b = g(o[1], wd), g(o[2], wd)
            mask = M64 if wd == 2 else M32
            if "_and_" in op:
                r = a & b
            elif "_or_" in op:
                r = a | b
            elif op.startswith("s_andn2"):
                r = a & ~b & mask
            else:
                r = (a | ~b) & mask
            w.sset(o[0], wd, r)
            w.scc = int(r != 0)
        elif op in ("s_cselect_b64", "s_cselect_b32"):
            wd = 2 if op.endswith("64") else 1
            w.sset(o[0], wd, g(o[1], wd) if w.scc else g(o[2], wd))
        elif op.startswith("s_cmp_"):
            kind = op[len("s_cmp_"):]
            if kind == "lg_u64":
                a, b = g(o[0], 2), g(o[1], 2)
                w.scc = int(a != b)
            else:
                a, b = g(o[0], 1), g(o[1], 1)
                cmp, ty = kind.split("_")
                if ty == "i32":
                    a, b = sx32(a), sx32(b)
                w.scc = int({"eq": a == b, "lg": a != b, "lt": a < b, "gt": a > b, "ge": a >= b, "le": a <= b}[cmp])
        elif op == "s_add_u32":
            r = g(o[1], 1) + g(o[2], 1)
            w.sset(o[0], 1, r & M32)
            w.scc = r >> 32
        elif op == "s_addc_u32":
            r = g(o[1], 1) + g(o[2], 1) + w.scc
            w.sset(o[0], 1, r & M32)
            w.scc = r >> 32
        elif op == "s_sub_u32":
            a, b = g(o[1], 1), g(o[2], 1)
            w.sset(o[0], 1, (a - b) & M32)
            w.scc = int(b > a)
        elif op == "s_subb_u32":
            a, b = g(o[1], 1), g(o[2], 1)
            w.sset(o[0], 1, (a - b - w.scc) & M32)
            w.scc = int(b + w.scc > a)
        elif op in ("s_add_i32", "s_sub_i32"):
            a, b = sx32(g(o[1], 1)), sx32(g(o[2], 1))
            r = a + b if op == "s_add_i32" else a - b
            w.sset(o[0], 1, r & M32)
            w.scc = int(r != sx32(r & M32))
        elif op == "s_mul_i32":
            w.sset(o[0], 1, (g(o[1], 1) * g(o[2], 1)) & M32)
        elif op == "s_mul_hi_u32":
            w.sset(o[0], 1, ((g(o[1], 1) * g(o[2], 1)) >> 32) & M32)
        elif op == "s_mul_hi_i32":
            w.sset(o[0], 1, ((sx32(g(o[1], 1)) * sx32(g(o[2], 1))) >> 32) & M32)
        elif op == "s_lshl_b32":
            r = (g(o[1], 1) << (g(o[2], 1) & 31)) & M32
            w.sset(o[0], 1, r)
            w.scc = int(r != 0)
        elif op == "s_lshl_b64":
            r = (g(o[1], 2) << (g(o[2], 1) & 63)) & M64
            w.sset(o[0], 2, r)
            w.scc = int(r != 0)
        elif op == "s_ashr_i32":
            r = (sx32(g(o[1], 1)) >> (g(o[2], 1) & 31)) & M32
            w.sset(o[0], 1, r)
            w.scc = int(r != 0)
        elif op == "s_min_i32":
            a, b = sx32(g(o[1], 1)), sx32(g(o[2], 1))
            w.sset(o[0], 1, min(a, b) & M32)
            w.scc = int(a <= b)
        elif op == "s_nor_b64":
            r = ~(g(o[1], 2) | g(o[2], 2)) & M64
            w.sset(o[0], 2, r)
            w.scc = int(r != 0)
        elif op == "s_abs_i32":
            r = abs(sx32(g(o[1], 1))) & M32
            w.sset(o[0], 1, r)
            w.scc = int(r != 0)
        elif op == "s_max_i32":
            a, b = sx32(g(o[1], 1)), sx32(g(o[2], 1))
            w.sset(o[0], 1, max(a, b) & M32)
            w.scc = int(a >= b)
        elif op == "s_xor_b32":
            r = g(o[1], 1) ^ g(o[2], 1)
            w.sset(o[0], 1, r)
            w.scc = int(r != 0)
        elif op == "s_xor_b64":
            r = g(o[1], 2) ^ g(o[2], 2)
            w.sset(o[0], 2, r)
            w.scc = int(r != 0)
        elif op == "s_movk_i32":
            w.sset(o[0], 1, sx16(int(o[1], 0)) & M32)
        elif op == "s_mulk_i32":
            w.sset(o[0], 1, (g(o[0], 1) * sx16(int(o[1], 0))) & M32)
        elif op == "s_lshr_b32":
            r = (g(o[1], 1) >> (g(o[2], 1) & 31)) & M32
            w.sset(o[0], 1, r)
            w.scc = int(r != 0)
        elif op == "s_brev_b32":
            w.sset(o[0], 1, int(format(g(o[1], 1), "032b")[::-1], 2))
        elif op in ("s_andn2_saveexec_b64", "s_or_saveexec_b64"):
            ex = g("exec", 2)
            src = g(o[1], 2)
            w.sset(o[0], 2, ex)
            ne = (src & ~ex & M64) if op == "s_andn2_saveexec_b64" else (src | ex)
            w.sset("exec", 2, ne)
            w.scc = int(ne != 0)
        elif op == "s_and_saveexec_b64":
            ex = g("exec", 2)
            src = g(o[1], 2)
            w.sset(o[0], 2, ex)
            ne = src & ex
            w.sset("exec", 2, ne)
            w.scc = int(ne != 0)
        else:
            raise NotImplementedError(op)
        return None

    def vmem(self, w, op, o, mods, pc):
        ar = w.ar
        act, _ = w.exec_mask()
        kind = op.split("_", 2)[-1]
        if kind in ("ushort", "short"):
            store = op.startswith("global_store")
            if store:
                vaddr, data, saddr = o[0], o[1], o[2]
            else:
                data, vaddr, saddr = o[0], o[1], o[2]
            off = mods.get("offset", 0)
            addr = w.vget(vaddr, 2) if saddr == "off" else np.full(NL, w.sget(saddr, 2), np.uint64) + w.vget(vaddr, 1)
            addr = addr + np.uint64(off & M64)
            a = addr[act]
            ar.check(a, 2, f"{op} pc={pc} {' '.join(o)}")
            if store:
                ar.scatter16(a, (w.vget(data, 1) & np.uint64(0xFFFF)).astype(np.uint16)[act])
            else:
                v = np.zeros(NL, np.uint64)
                v[act] = ar.gather16(a)
                w.vset(data, 1, v, act)
            return None
        n = {"dword": 1, "dwordx2": 2, "dwordx3": 3, "dwordx4": 4}[op.split("_")[-1]]
        store = op.startswith("global_store")
        if store:
            vaddr, data, saddr = o[0], o[1], o[2]
        else:
            data, vaddr, saddr = o[0], o[1], o[2]
        off = mods.get("offset", 0)
        if saddr == "off":
            addr = w.vget(vaddr, 2)
        else:
            addr = np.full(NL, w.sget(saddr, 2), np.uint64) + w.vget(vaddr, 1)
        addr = addr + np.uint64(off & M64)
        a = addr[act]
        ar.check(a, 4 * n, f"{op} pc={pc} {' '.join(o)}")
        m = re.fullmatch(r"([va])\[(\d+):(\d+)\]", data) or re.fullmatch(r"([va])(\d+)()", data)
        R = w.V if m.group(1) == "v" else w.A
        r0 = int(m.group(2))
        if store:
            for k in range(n):
                ar.scatter32(a + np.uint64(4 * k), R[r0 + k][act])
        else:
            for k in range(n):
                R[r0 + k][act] = ar.gather32(a + np.uint64(4 * k))
        return None

    def vatomic(self, w, op, o, mods, pc):
        ar = w.ar
        act, _ = w.exec_mask()
        ret = mods.get("glc", False) or mods.get("sc0", False)
        if ret:
            dst, vaddr, data, saddr = o[0], o[1], o[2], o[3]
        else:
            vaddr, data, saddr = o[0], o[1], o[2]
        f64op = op.endswith("f64")
        n = 2 if f64op else 1
        off = mods.get("offset", 0)
        addr = w.vget(vaddr, 2) if saddr == "off" else np.full(NL, w.sget(saddr, 2), np.uint64) + w.vget(vaddr, 1)
        addr = addr + np.uint64(off & M64)
        ar.check(addr[act], 4 * n, f"{op} pc={pc}")
        val = w.vget(data, n)
        if not ret:
            a = addr[act]
            off = a.astype(np.int64) - ar.BASE
            if n == 2:
                assert (off % 8 == 0).all()
                np.add.at(ar.mem.view(np.float64), off // 8, f64(val[act]))
            else:
                np.add.at(ar.mem.view(np.float32), off // 4, f32(val[act]))
            return None
        old = np.zeros(NL, np.uint64)
        for l in np.nonzero(act)[0]:
            a = int(addr[l])
            cur = ar.gather32(np.array([a], np.uint64)).astype(np.uint64)
            if n == 2:
                cur = cur | (ar.gather32(np.array([a + 4], np.uint64)).astype(np.uint64) << np.uint64(32))
                nv = u64(f64(cur) + f64(np.array([val[l]], np.uint64)))
                ar.scatter32(np.array([a], np.uint64), (nv & np.uint64(M32)).astype(np.uint32))
                ar.scatter32(np.array([a + 4], np.uint64), (nv >> np.uint64(32)).astype(np.uint32))
            else:
                nv = u32f(f32(cur) + f32(np.array([val[l]], np.uint64)))
                ar.scatter32(np.array([a], np.uint64), nv.astype(np.uint32))
            old[l] = cur[0]
        if ret:
            w.vset(dst, n, old, act)
        return None

    def lds(self, w, op, o, mods, pc):
        act, _ = w.exec_mask()
        mem = w.lds.view(np.uint32)

        def chk(a, nbytes):
            a = a[act]
            if a.size and ((a + nbytes > w.lds_size).any() or (a % 4 != 0).any()):
                raise OOB(f"{op} pc={pc}: LDS address out of bounds ({int(a.max())} + {nbytes} > {w.lds_size})")

        base = w.vget(o[1] if op.startswith("ds_read") else o[0], 1).astype(np.int64)
        if op in ("ds_write_b32", "ds_read_b32"):
            a = base + mods.get("offset", 0)
            chk(a, 4)
            if op == "ds_write_b32":
                mem[(a // 4)[act]] = (w.vget(o[1], 1) & np.uint64(M32)).astype(np.uint32)[act]
            else:
                v = np.zeros(NL, np.uint64)
                v[act] = mem[(a // 4)[act]]
                w.vset(o[0], 1, v, act)
            return None
        if op == "ds_read2_b32":
            a0 = base + mods.get("offset0", 0) * 4
            a1 = base + mods.get("offset1", 0) * 4
            chk(a0, 4)
            chk(a1, 4)
            m = re.fullmatch(r"v\[(\d+):(\d+)\]", o[0])
            r0 = int(m.group(1))
            for q, a in enumerate((a0, a1)):
                v = np.zeros(NL, np.uint64)
                v[act] = mem[(a // 4)[act]]
                w.vset(f"v{r0 + q}", 1, v, act)
            return None
        if op in ("ds_write_b64", "ds_read_b64"):
            a = base + mods.get("offset", 0)
            chk(a, 8)
            if op == "ds_write_b64":
                v = w.vget(o[1], 2)
                mem[(a // 4)[act]] = (v & np.uint64(M32)).astype(np.uint32)[act]
                mem[(a // 4 + 1)[act]] = (v >> np.uint64(32)).astype(np.uint32)[act]
            else:
                lo = np.zeros(NL, np.uint64)
                hi = np.zeros(NL, np.uint64)
                lo[act] = mem[(a // 4)[act]]
                hi[act] = mem[(a // 4 + 1)[act]]
                w.vset(o[0], 2, lo | (hi << np.uint64(32)), act)
        elif op in ("ds_write2st64_b64", "ds_read2st64_b64"):
            a0 = base + mods.get("offset0", 0) * 64 * 8
            a1 = base + mods.get("offset1", 0) * 64 * 8
            chk(a0, 8)
            chk(a1, 8)
            if op == "ds_write2st64_b64":
                for a, src in ((a0, o[1]), (a1, o[2])):
                    v = w.vget(src, 2)
                    mem[(a // 4)[act]] = (v & np.uint64(M32)).astype(np.uint32)[act]
                    mem[(a // 4 + 1)[act]] = (v >> np.uint64(32)).astype(np.uint32)[act]
            else:
                m = re.fullmatch(r"v\[(\d+):(\d+)\]", o[0])
                r0 = int(m.group(1))
                for q, a in enumerate((a0, a1)):
                    lo = np.zeros(NL, np.uint64)
                    hi = np.zeros(NL, np.uint64)
                    lo[act] = mem[(a // 4)[act]]
                    hi[act] = mem[(a // 4 + 1)[act]]
                    w.vset(f"v[{r0 + 2 * q}:{r0 + 2 * q + 1}]", 2, lo | (hi << np.uint64(32)), act)
        else:
            raise NotImplementedError(op)
        return None

    def sdwa(self, w, base, o, mods):
        """SDWA form: operands may be sub-dword selections (BYTE_n / WORD_n),
        optionally sign-extended (sext(...)); the result is written as a full
        dword (dst_sel DWORD, unused bits zero)."""
        act, _ = w.exec_mask()
        assert mods.get("dst_sel", "DWORD") == "DWORD", mods

        def src(k, sel):
            t = o[k]
            sx = t.startswith("sext(")
            if sx:
                t = t[5:-1]
            u = w.vget(t, 1)
            if sel.startswith("BYTE_"):
                sh, bits = 8 * int(sel[5]), 8
            elif sel.startswith("WORD_"):
                sh, bits = 16 * int(sel[5]), 16
            else:
                return u
            x = (u >> np.uint64(sh)) & np.uint64((1 << bits) - 1)
            if sx:
                x = np.where(x >> np.uint64(bits - 1), x | np.uint64(M32 ^ ((1 << bits) - 1)), x)
            return x & np.uint64(M32)

        a = src(1, mods.get("src0_sel", "DWORD"))
        if base == "v_cvt_f32_f16":  # unary: half in the selected word
            h = (a & np.uint64(0xFFFF)).astype(np.uint16).view(np.float16)
            w.vset(o[0], 1, u32f(h.astype(np.float32)), act)
            return None
        b = src(2, mods.get("src1_sel", "DWORD"))
        r = self.int_op(base, a, b)
        w.vset(o[0], 1, r, act)
        return None

    @staticmethod
    def int_op(base, a, b, c=None):
        """Integer VALU operations on uint64 lane arrays holding 32-bit values."""
        M = np.uint64(M32)
        s24 = lambda x: ((x & np.uint64(0xFFFFFF)).astype(np.int64) ^ 0x800000) - 0x800000
        if base == "v_mul_i32_i24":
            return (s24(a) * s24(b)).astype(np.uint64) & M
        if base == "v_mul_u32_u24":
            return ((a & np.uint64(0xFFFFFF)) * (b & np.uint64(0xFFFFFF))) & M
        if base == "v_add_u16":
            return (a + b) & np.uint64(0xFFFF)
        if base == "v_mul_lo_u16":
            return ((a & np.uint64(0xFFFF)) * (b & np.uint64(0xFFFF))) & np.uint64(0xFFFF)
        if base == "v_ashrrev_i32":
            return (s32(b) >> (a & np.uint64(31)).astype(np.int64)).astype(np.uint64) & M
        raise NotImplementedError(base)

    def valu(self, w, op, o):
        act, _ = w.exec_mask()
        V = w.vget
        if op == "v_mov_b32_e32":
            w.vset(o[0], 1, V(o[1], 1), act)
        elif op == "v_mov_b64_e32":
            w.vset(o[0], 2, V(o[1], 2, True), act)
        elif op in ("v_accvgpr_read_b32", "v_accvgpr_write_b32", "v_accvgpr_mov_b32"):
            w.vset(o[0], 1, V(o[1], 1), act)
        elif op == "v_readlane_b32":
            lane = int(o[2]) if o[2].isdigit() else w.sget(o[2], 1) & 63
            w.sset(o[0], 1, int(V(o[1], 1)[lane]))
        elif op == "v_writelane_b32":
            lane = int(o[2]) if o[2].isdigit() else w.sget(o[2], 1) & 63
            m = np.zeros(NL, bool)
            m[lane] = True
            w.vset(o[0], 1, np.full(NL, w.sget(o[1], 1), np.uint64), m)
        elif op == "v_add_u32_e32":
            w.vset(o[0], 1, (V(o[1], 1) + V(o[2], 1)) & np.uint64(M32), act)
        elif op == "v_subrev_u32_e32":
            w.vset(o[0], 1, (V(o[2], 1) - V(o[1], 1)) & np.uint64(M32), act)
        elif op == "v_add3_u32":
            w.vset(o[0], 1, (V(o[1], 1) + V(o[2], 1) + V(o[3], 1)) & np.uint64(M32), act)
        elif op == "v_mul_lo_u32":
            w.vset(o[0], 1, (V(o[1], 1) * V(o[2], 1)) & np.uint64(M32), act)
        elif op == "v_ashrrev_i32_e32":
            sh = V(o[1], 1) & np.uint64(31)
            w.vset(o[0], 1, (s32(V(o[2], 1)) >> sh.astype(np.int64)).astype(np.uint64) & np.uint64(M32), act)
        elif op == "v_lshrrev_b32_e32":
            w.vset(o[0], 1, (V(o[2], 1) >> (V(o[1], 1) & np.uint64(31))), act)
        elif op == "v_lshlrev_b64":
            w.vset(o[0], 2, V(o[2], 2) << (V(o[1], 1) & np.uint64(63)), act)
        elif op == "v_lshl_or_b32":
            w.vset(o[0], 1, ((V(o[1], 1) << (V(o[2], 1) & np.uint64(31))) | V(o[3], 1)) & np.uint64(M32), act)
        elif op == "v_and_or_b32":
            w.vset(o[0], 1, (V(o[1], 1) & V(o[2], 1)) | V(o[3], 1), act)
        elif op == "v_perm_b32":
            # byte select from {src0 (bytes 4..7), src1 (bytes 0..3)}; 8..11 sign
            # fills, 12 zero, 13.. 0xff (CDNA ISA, V_PERM_B32)
            s0, s1, sel = V(o[1], 1), V(o[2], 1), V(o[3], 1)
            comb = (s0 << np.uint64(32)) | s1
            res = np.zeros(NL, np.uint64)
            for b in range(4):
                sb = (sel >> np.uint64(8 * b)) & np.uint64(0xFF)
                byte = (comb >> (np.uint64(8) * np.minimum(sb, np.uint64(7)))) & np.uint64(0xFF)
                ff = np.uint64(0xFF)
                byte = np.where(sb == 8, ((s1 >> np.uint64(15)) & np.uint64(1)) * ff, byte)
                byte = np.where(sb == 9, ((s1 >> np.uint64(31)) & np.uint64(1)) * ff, byte)
                byte = np.where(sb == 10, ((s0 >> np.uint64(15)) & np.uint64(1)) * ff, byte)
                byte = np.where(sb == 11, ((s0 >> np.uint64(31)) & np.uint64(1)) * ff, byte)
                byte = np.where(sb == 12, np.uint64(0), byte)
                byte = np.where(sb >= 13, ff, byte)
                res |= byte << np.uint64(8 * b)
            w.vset(o[0], 1, res, act)
        elif op == "v_not_b32_e32":
            w.vset(o[0], 1, ~V(o[1], 1) & np.uint64(M32), act)
        elif op == "v_lshl_add_u64":
            w.vset(o[0], 2, (V(o[1], 2) << (V(o[2], 1) & np.uint64(63))) + V(o[3], 2), act)
        elif op in ("v_mad_u64_u32", "v_mad_i64_i32"):
            a, b, c = V(o[2], 1), V(o[3], 1), V(o[4], 2)
            if op == "v_mad_u64_u32":
                lo = [((int(a[l]) * int(b[l]) + int(c[l])) & M64) for l in range(NL)]
            else:
                lo = [((sx32(int(a[l])) * sx32(int(b[l])) + (int(c[l]) - (1 << 64) if int(c[l]) >> 63 else int(c[l]))) & M64)
                      for l in range(NL)]
            w.vset(o[0], 2, np.array(lo, np.uint64), act)
            # carry/overflow output is not used for addressing by these kernels
            w.sset(o[1], 2, 0)
        elif op == "v_subrev_co_u32_e32":
            a, b = V(o[2], 1), V(o[3], 1)
            w.vset(o[0], 1, (b - a) & np.uint64(M32), act)
            w.sset(o[1], 2, w.bits_of(act & (a > b)))
        elif op == "v_subb_co_u32_e32":
            a, b = V(o[2], 1), V(o[3], 1)
            cin = w.mask_of(w.sget(o[4], 2)).astype(np.uint64)
            w.vset(o[0], 1, (a - b - cin) & np.uint64(M32), act)
            w.sset(o[1], 2, w.bits_of(act & (b + cin > a)))
        elif op == "v_cndmask_b32_e32" or op == "v_cndmask_b32_e64":
            msk = w.mask_of(w.sget(o[3] if len(o) > 3 else "vcc", 2))
            w.vset(o[0], 1, np.where(msk, V(o[2], 1), V(o[1], 1)), act)
        elif op.startswith("v_cmp_"):
            kind = op[len("v_cmp_"):].rsplit("_", 1)[0]
            dst = o[0]
            cmp, ty = kind.split("_")
            wd = 2 if ty in ("f64", "i64", "u64") else 1
            if cmp == "class":
                a = V(o[1], wd, True)
                x = f64(a) if wd == 2 else f32(a)
                bits = V(o[2], 1).astype(np.int64)
                neg = np.signbit(x)
                tiny = np.finfo(np.float64 if wd == 2 else np.float32).tiny
                with np.errstate(invalid="ignore"):
                    cls = np.where(np.isnan(x), 1, np.where(np.isinf(x), np.where(neg, 2, 9),
                                   np.where(x == 0, np.where(neg, 5, 6),
                                            np.where(np.abs(x) < tiny, np.where(neg, 4, 7), np.where(neg, 3, 8)))))
                r = ((bits >> cls) & 1).astype(bool)
            else:
                a, b = V(o[1], wd, ty.startswith("f")), V(o[2], wd, ty.startswith("f"))
                if ty == "i32":
                    a, b = s32(a), s32(b)
                elif ty == "i64":
                    a, b = a.view(np.int64), b.view(np.int64)
                elif ty == "f32":
                    a, b = f32(a), f32(b)
                elif ty == "f64":
                    a, b = f64(a), f64(b)
                with np.errstate(invalid="ignore"):
                    r = {"eq": lambda: a == b, "ne": lambda: a != b, "lt": lambda: a < b, "gt": lambda: a > b,
                         "ge": lambda: a >= b, "le": lambda: a <= b, "lg": lambda: (a < b) | (a > b),
                         "neq": lambda: ~(a == b), "ngt": lambda: ~(a > b), "nlt": lambda: ~(a < b),
                         "nge": lambda: ~(a >= b), "nle": lambda: ~(a <= b), "nlg": lambda: ~((a < b) | (a > b)),
                         "o": lambda: ~(np.isnan(a) | np.isnan(b)), "u": lambda: np.isnan(a) | np.isnan(b)}[cmp]()
            w.sset(dst, 2, w.bits_of(act & r))
        elif op in ("v_mul_f32_e32", "v_fma_f32", "v_fmac_f32_e32", "v_sqrt_f32_e32", "v_cvt_f32_f64_e32"):
            if op == "v_mul_f32_e32":
                r = f32(V(o[1], 1)) * f32(V(o[2], 1))
            elif op == "v_fma_f32":
                r = (f32(V(o[1], 1)).astype(np.float64) * f32(V(o[2], 1)) + f32(V(o[3], 1))).astype(np.float32)
            elif op == "v_fmac_f32_e32":
                r = (f32(V(o[1], 1)).astype(np.float64) * f32(V(o[2], 1)) + f32(V(o[0], 1))).astype(np.float32)
            elif op == "v_sqrt_f32_e32":
                with np.errstate(invalid="ignore"):
                    r = np.sqrt(f32(V(o[1], 1)))
            else:
                r = f64(V(o[1], 2, True)).astype(np.float32)
            w.vset(o[0], 1, u32f(r), act)
        elif op in ("v_fma_f64", "v_mul_f64", "v_add_f64", "v_fmac_f64_e32", "v_rcp_f64_e32", "v_div_fmas_f64",
                    "v_div_fixup_f64"):
            with np.errstate(all="ignore"):
                if op == "v_fma_f64" or op == "v_div_fmas_f64":
                    r = f64(V(o[1], 2, True)) * f64(V(o[2], 2, True)) + f64(V(o[3], 2, True))
                elif op == "v_mul_f64":
                    r = f64(V(o[1], 2, True)) * f64(V(o[2], 2, True))
                elif op == "v_add_f64":
                    r = f64(V(o[1], 2, True)) + f64(V(o[2], 2, True))
                elif op == "v_fmac_f64_e32":
                    r = f64(V(o[1], 2, True)) * f64(V(o[2], 2, True)) + f64(V(o[0], 2, True))
                elif op == "v_rcp_f64_e32":
                    r = 1.0 / f64(V(o[1], 2, True))
                else:  # div_fixup: quotient for finite normal inputs
                    r = f64(V(o[1], 2, True))
            w.vset(o[0], 2, u64(r), act)
        elif op == "v_div_scale_f64":
            w.vset(o[0], 2, V(o[2], 2, True), act)
            w.sset(o[1], 2, 0)
        elif op in ("v_add_f32_e32", "v_sub_f32_e32", "v_sub_f32_e64", "v_mul_f32_e64", "v_fmamk_f32", "v_fmaak_f32",
                    "v_rcp_f32_e32", "v_rcp_iflag_f32_e32", "v_exp_f32_e32", "v_trunc_f32_e32", "v_rndne_f32_e32",
                    "v_div_fmas_f32", "v_ldexp_f32", "v_cvt_f32_u32_e32", "v_cvt_f32_i32_e32", "v_cvt_f32_f64_e64"):
            with np.errstate(all="ignore"):
                x = lambda k: f32(V(o[k], 1)).astype(np.float64)
                if op == "v_add_f32_e32":
                    r = x(1) + x(2)
                elif op in ("v_sub_f32_e32", "v_sub_f32_e64"):
                    r = x(1) - x(2)
                elif op == "v_mul_f32_e64":
                    r = x(1) * x(2)
                elif op in ("v_fmamk_f32", "v_fmaak_f32"):
                    # fmamk: src0 * K + src1, fmaak: src0 * src1 + K (operands in that order)
                    r = x(1) * x(2) + x(3)
                elif op in ("v_rcp_f32_e32", "v_rcp_iflag_f32_e32"):
                    r = 1.0 / x(1)
                elif op == "v_exp_f32_e32":
                    r = np.exp2(x(1))
                elif op == "v_trunc_f32_e32":
                    r = np.trunc(x(1))
                elif op == "v_rndne_f32_e32":
                    r = np.rint(x(1))
                elif op == "v_div_fmas_f32":
                    r = x(1) * x(2) + x(3)
                elif op == "v_ldexp_f32":
                    r = np.ldexp(x(1), s32(V(o[2], 1)))
                elif op == "v_cvt_f32_u32_e32":
                    r = V(o[1], 1).astype(np.float64)
                elif op == "v_cvt_f32_i32_e32":
                    r = s32(V(o[1], 1)).astype(np.float64)
                else:
                    r = f64(V(o[1], 2, True))
            w.vset(o[0], 1, u32f(r.astype(np.float32)), act)
        elif op == "v_div_scale_f32":
            w.vset(o[0], 1, V(o[2], 1), act)
            w.sset(o[1], 2, 0)
        elif op == "v_div_fixup_f32":
            w.vset(o[0], 1, V(o[1], 1), act)
        elif op in ("v_rsq_f64_e32", "v_ldexp_f64", "v_max_f64", "v_min_f64", "v_rndne_f64_e32", "v_cvt_f64_f32_e32",
                    "v_cvt_f64_i32_e32"):
            with np.errstate(all="ignore"):
                if op == "v_rsq_f64_e32":
                    r = 1.0 / np.sqrt(f64(V(o[1], 2, True)))
                elif op == "v_ldexp_f64":
                    r = np.ldexp(f64(V(o[1], 2, True)), s32(V(o[2], 1)))
                elif op == "v_max_f64":
                    r = np.fmax(f64(V(o[1], 2, True)), f64(V(o[2], 2, True)))
                elif op == "v_min_f64":
                    r = np.fmin(f64(V(o[1], 2, True)), f64(V(o[2], 2, True)))
                elif op == "v_rndne_f64_e32":
                    r = np.rint(f64(V(o[1], 2, True)))
                elif op == "v_cvt_f64_f32_e32":
                    r = f32(V(o[1], 1)).astype(np.float64)
                else:
                    r = s32(V(o[1], 1)).astype(np.float64)
            w.vset(o[0], 2, u64(r), act)
        elif op in ("v_cvt_u32_f32_e32", "v_cvt_i32_f32_e32", "v_cvt_i32_f64_e32"):
            with np.errstate(all="ignore"):
                x = f64(V(o[1], 2, True)) if op.endswith("f64_e32") else f32(V(o[1], 1)).astype(np.float64)
                x = np.nan_to_num(np.trunc(x), nan=0.0)
                if op == "v_cvt_u32_f32_e32":
                    r = np.clip(x, 0, 4294967295.0).astype(np.uint64)
                else:
                    r = np.clip(x, -2147483648.0, 2147483647.0).astype(np.int64).astype(np.uint64) & np.uint64(M32)
            w.vset(o[0], 1, r, act)
        elif op == "v_cvt_pk_bf16_f32":
            def bf(u):
                u = u & np.uint64(M32)
                return ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) & np.uint64(0xFFFF)
            w.vset(o[0], 1, bf(V(o[1], 1)) | (bf(V(o[2], 1)) << np.uint64(16)), act)
        elif op in ("v_xor_b32_e32", "v_or_b32_e32", "v_and_b32_e32"):
            a, b = V(o[1], 1), V(o[2], 1)
            w.vset(o[0], 1, {"x": a ^ b, "o": a | b, "a": a & b}[op[2]], act)
        elif op == "v_lshlrev_b32_e32":
            w.vset(o[0], 1, (V(o[2], 1) << (V(o[1], 1) & np.uint64(31))) & np.uint64(M32), act)
        elif op == "v_mul_hi_u32":
            w.vset(o[0], 1, (V(o[1], 1) * V(o[2], 1)) >> np.uint64(32), act)
        elif op == "v_med3_i32":
            x = np.stack([s32(V(o[k], 1)) for k in (1, 2, 3)])
            w.vset(o[0], 1, np.median(x, axis=0).astype(np.int64).astype(np.uint64) & np.uint64(M32), act)
        elif op == "v_bfe_u32":
            off, wid = V(o[2], 1) & np.uint64(31), V(o[3], 1) & np.uint64(31)
            w.vset(o[0], 1, (V(o[1], 1) >> off) & ((np.uint64(1) << wid) - np.uint64(1)), act)
        elif op == "v_sub_u32_e32":
            w.vset(o[0], 1, (V(o[1], 1) - V(o[2], 1)) & np.uint64(M32), act)
        elif op in ("v_min_i32_e32", "v_max_i32_e32"):
            f = np.minimum if op == "v_min_i32_e32" else np.maximum
            w.vset(o[0], 1, f(s32(V(o[1], 1)), s32(V(o[2], 1))).astype(np.int64).astype(np.uint64) & np.uint64(M32), act)
        elif op in ("v_sub_co_u32_e32", "v_subrev_co_u32_e64", "v_add_co_u32_e32"):
            a, b = V(o[2], 1), V(o[3], 1)
            if op == "v_subrev_co_u32_e64":
                a, b = b, a
            if op == "v_add_co_u32_e32":
                r = a + b
                w.vset(o[0], 1, r & np.uint64(M32), act)
                w.sset(o[1], 2, w.bits_of(act & (r > np.uint64(M32))))
            else:
                w.vset(o[0], 1, (a - b) & np.uint64(M32), act)
                w.sset(o[1], 2, w.bits_of(act & (b > a)))
        elif op in ("v_subb_co_u32_e64", "v_subbrev_co_u32_e64", "v_addc_co_u32_e32", "v_addc_co_u32_e64"):
            a, b = V(o[2], 1), V(o[3], 1)
            cin = w.mask_of(w.sget(o[4], 2)).astype(np.uint64)
            if op == "v_subbrev_co_u32_e64":
                a, b = b, a
            if op.startswith("v_addc"):
                r = a + b + cin
                w.vset(o[0], 1, r & np.uint64(M32), act)
                w.sset(o[1], 2, w.bits_of(act & (r > np.uint64(M32))))
            else:
                w.vset(o[0], 1, (a - b - cin) & np.uint64(M32), act)
                w.sset(o[1], 2, w.bits_of(act & (b + cin > a)))
        elif op == "v_cvt_f32_f16_e32":
            h = (V(o[1], 1) & np.uint64(0xFFFF)).astype(np.uint16).view(np.float16)
            w.vset(o[0], 1, u32f(h.astype(np.float32)), act)
        elif op in ("v_mul_i32_i24_e32", "v_mul_u32_u24_e32", "v_add_u16_e32", "v_mul_lo_u16_e32"):
            w.vset(o[0], 1, self.int_op(op[:-4], V(o[1], 1), V(o[2], 1)), act)
        elif op == "v_mad_i32_i24":
            s24 = lambda x: ((x & np.uint64(0xFFFFFF)).astype(np.int64) ^ 0x800000) - 0x800000
            w.vset(o[0], 1, (s24(V(o[1], 1)) * s24(V(o[2], 1)) + s32(V(o[3], 1))).astype(np.uint64) & np.uint64(M32), act)
        elif op == "v_bfe_i32":
            off, wid = V(o[2], 1) & np.uint64(31), V(o[3], 1) & np.uint64(31)
            x = (V(o[1], 1) >> off) & ((np.uint64(1) << wid) - np.uint64(1))
            sign = (x >> (wid - np.uint64(1))) & np.uint64(1)
            x = np.where((wid > 0) & (sign == 1), x | (np.uint64(M32) ^ ((np.uint64(1) << wid) - np.uint64(1))), x)
            w.vset(o[0], 1, x & np.uint64(M32), act)
        elif op == "v_lshl_add_u32":
            w.vset(o[0], 1, ((V(o[1], 1) << (V(o[2], 1) & np.uint64(31))) + V(o[3], 1)) & np.uint64(M32), act)
        elif op == "v_add_lshl_u32":
            w.vset(o[0], 1, ((V(o[1], 1) + V(o[2], 1)) << (V(o[3], 1) & np.uint64(31))) & np.uint64(M32), act)
        elif op == "v_mul_hi_i32":
            w.vset(o[0], 1, ((s32(V(o[1], 1)) * s32(V(o[2], 1))) >> 32).astype(np.uint64) & np.uint64(M32), act)
        elif op in ("v_ashrrev_i16_e32", "v_lshrrev_b16_e32"):
            sh = V(o[1], 1) & np.uint64(15)
            x = V(o[2], 1) & np.uint64(0xFFFF)
            if op == "v_ashrrev_i16_e32":
                xs = (x.astype(np.int64) ^ 0x8000) - 0x8000
                r = (xs >> sh.astype(np.int64)).astype(np.uint64) & np.uint64(0xFFFF)
            else:
                r = x >> sh
            w.vset(o[0], 1, r, act)
        elif op == "v_readfirstlane_b32":
            act2 = np.nonzero(act)[0]
            w.sset(o[0], 1, int(V(o[1], 1)[act2[0] if act2.size else 0]))
        else:
            raise NotImplementedError(op)
        return None
