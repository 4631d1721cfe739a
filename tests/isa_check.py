"""Lane-aware check of the address operands of memory instructions in the AMD assembly
(gfx950, wave64) of the code objects built here.

Why: a vector register is written only in the lanes active (EXEC) when the writing
instruction runs.  If the compiler merges per-path stores into one store after a divergent
join, lanes that took a path which never wrote the address register still execute the store
through whatever that register held: a GPU memory fault.  This once happened in the cache
kernel (a merged store of a pair's stage whose address sphere/capsule lanes never wrote).

What: for each kernel, a forward must-analysis over the machine control-flow graph tracks
EXEC, the 64-bit lane masks in SGPR pairs and VCC as boolean functions (a small reduced
ordered BDD) over atoms (one atom per comparison or other unknown mask source, stable per
instruction), and for each VGPR the set of lanes in which it has surely been written.  At
every global/flat/buffer load, store or atomic, the lanes in EXEC must be lanes in which
each address VGPR has been written.  An s_cbranch_execz/execnz edge taken because EXEC is
empty makes the skipped region's lanes vacuous on that edge.  Joins take the union of EXEC
and of each mask and the intersection of the written-lane sets; a mask known on only some
predecessors becomes a fresh atom.  Comparisons inside loops reuse their atom on every iteration, and
partial writes (v_writelane) do not count as writes.

Used by tests/test_isa_store_addresses.py only."""
import collections
import re

# ----------------------------------------------------------------------------- BDD


class BDD:
    """Reduced ordered BDD: node 0 = false, 1 = true, others (var, lo, hi)."""

    def __init__(self):
        self.nodes = [None, None]
        self.unique = {}
        self.memo = {}
        self.nvars = 0

    def var(self):
        v = self.nvars
        self.nvars += 1
        return self.mk(v, 0, 1)

    def mk(self, v, lo, hi):
        if lo == hi:
            return lo
        key = (v, lo, hi)
        n = self.unique.get(key)
        if n is None:
            n = len(self.nodes)
            self.nodes.append(key)
            self.unique[key] = n
        return n

    def top(self, u):
        return self.nodes[u][0] if u > 1 else 1 << 60

    def ite(self, f, g, h):
        if f == 1:
            return g
        if f == 0:
            return h
        if g == h:
            return g
        if g == 1 and h == 0:
            return f
        key = (f, g, h)
        r = self.memo.get(key)
        if r is not None:
            return r
        v = min(self.top(f), self.top(g), self.top(h))

        def cof(u, b):
            if u <= 1 or self.nodes[u][0] != v:
                return u
            return self.nodes[u][2] if b else self.nodes[u][1]

        r = self.mk(v, self.ite(cof(f, 0), cof(g, 0), cof(h, 0)), self.ite(cof(f, 1), cof(g, 1), cof(h, 1)))
        self.memo[key] = r
        return r

    def AND(self, a, b):
        return self.ite(a, b, 0)

    def OR(self, a, b):
        return self.ite(a, 1, b)

    def NOT(self, a):
        return self.ite(a, 0, 1)

    def XOR(self, a, b):
        return self.ite(a, self.NOT(b), b)


# ----------------------------------------------------------------------------- parsing

LABEL = re.compile(r"^([.\w$]+):")
VREG = re.compile(r"^v\[(\d+):(\d+)\]$|^v(\d+)$")
SPAIR = re.compile(r"^s\[(\d+):(\d+)\]$")


def vregs(tok):
    m = VREG.match(tok.strip())
    if not m:
        return []
    if m.group(3) is not None:
        return [int(m.group(3))]
    return list(range(int(m.group(1)), int(m.group(2)) + 1))


def sregs(tok):
    tok = tok.strip()
    m = SPAIR.match(tok)
    if m:
        return list(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"^s(\d+)$", tok)
    return [int(m.group(1))] if m else []


def kernels(text):
    """(name, [lines]) of each function body in a device assembly file."""
    cur = None
    for line in text.splitlines():
        m = re.match(r"^([A-Za-z_][\w.$]*):\s*(;.*)?$", line)
        if m:
            if cur:
                yield cur
            cur = (m.group(1), [])
            continue
        if cur is not None:
            if line.strip().startswith(".Lfunc_end"):
                yield cur
                cur = None
                continue
            cur[1].append(line)
    if cur:
        yield cur


def instructions(lines):
    """[(label or None, opcode, operands)] with comments and directives dropped."""
    out = []
    for raw in lines:
        line = raw.split(";")[0].strip()
        if not line:
            continue
        m = LABEL.match(line)
        if m:
            out.append((m.group(1), None, []))
            continue
        if line.startswith("."):
            continue
        parts = line.split(None, 1)
        op = parts[0]
        ops = [o.strip() for o in parts[1].split(",")] if len(parts) > 1 else []
        # the last operand may carry modifiers after a space ("off offset:16", "s[4:7], 0 offen")
        ops = [o.split()[0] if o else o for o in ops]
        mods = parts[1] if len(parts) > 1 else ""
        out.append((None, op, (ops, mods)))
    return out


# ----------------------------------------------------------------------------- analysis

MEM = re.compile(r"^(global|flat|buffer)_(load|store|atomic)")


def address_operands(op, ops, mods):
    """The address VGPRs of a global/flat/buffer access."""
    kind = MEM.match(op)
    space, what = kind.group(1), kind.group(2)
    if space == "buffer":
        if "offen" not in mods and "idxen" not in mods:
            return []
        i = 1
        return vregs(ops[i]) if len(ops) > i else []
    if what == "load":
        return vregs(ops[1]) if len(ops) > 1 else []
    if what == "store":
        return vregs(ops[0])
    # atomics: returning forms carry a destination first (glc / sc0)
    n = len(ops)
    ret = ("glc" in mods or "sc0" in mods) and n >= (4 if space == "global" else 3)
    return vregs(ops[1] if ret else ops[0])


def access_dests(op, ops, mods):
    kind = MEM.match(op)
    if kind.group(2) == "load":
        return vregs(ops[0])
    if kind.group(2) == "atomic" and ("glc" in mods or "sc0" in mods):
        return vregs(ops[0])
    return []


class State:
    __slots__ = ("exec", "masks", "defs")

    def __init__(self, exec_, masks, defs):
        self.exec = exec_
        self.masks = masks
        self.defs = defs

    def copy(self):
        return State(self.exec, dict(self.masks), dict(self.defs))

    def key(self):
        return (self.exec, tuple(sorted(self.masks.items())),
                tuple(sorted(self.defs.items(), key=lambda kv: str(kv[0]))))


class Kernel:
    def __init__(self, name, lines, bdd):
        self.name = name
        self.bdd = bdd
        self.ins = instructions(lines)
        self.atoms = {}
        self.findings = []
        self._blocks()

    def atom(self, key):
        a = self.atoms.get(key)
        if a is None:
            a = self.atoms[key] = self.bdd.var()
        return a

    def _blocks(self):
        starts = {0}
        labels = {}
        for i, (lab, op, _) in enumerate(self.ins):
            if lab is not None:
                starts.add(i)
                labels[lab] = i
            elif op and (op.startswith("s_branch") or op.startswith("s_cbranch") or op == "s_endpgm"):
                starts.add(i + 1)
        self.starts = sorted(s for s in starts if s < len(self.ins))
        self.labels = labels
        self.block_of = {}
        for k, s in enumerate(self.starts):
            e = self.starts[k + 1] if k + 1 < len(self.starts) else len(self.ins)
            self.block_of[s] = e

    def mask_value(self, st, tok, site):
        b = self.bdd
        tok = tok.strip()
        if tok == "exec":
            return st.exec
        if tok in ("vcc", "exec_lo", "vcc_lo"):
            return st.masks.get("vcc", self.atom(("unknown", site, tok)))
        if tok in ("0", "0x0"):
            return 0
        if tok in ("-1", "0xffffffffffffffff"):
            return 1
        if SPAIR.match(tok):
            return st.masks.get(tok, self.atom(("unknown", site, tok)))
        return self.atom(("unknown", site, tok))

    def set_mask(self, st, tok, val):
        tok = tok.strip()
        if tok == "exec":
            st.exec = val
        elif tok == "vcc":
            st.masks["vcc"] = val
        elif SPAIR.match(tok):
            self.kill_sregs(st, sregs(tok))
            st.masks[tok] = val

    def kill_sregs(self, st, regs):
        if not regs:
            return
        rs = set(regs)
        for k in list(st.masks):
            if k != "vcc" and rs & set(sregs(k)):
                del st.masks[k]

    def step(self, st, i, check):
        lab, op, payload = self.ins[i]
        if lab is not None:
            return
        ops, mods = payload
        b = self.bdd
        site = i
        if op.startswith("s_") and op.endswith("saveexec_b64"):
            src = self.mask_value(st, ops[1], site)
            old = st.exec
            self.set_mask(st, ops[0], old)
            if op == "s_and_saveexec_b64":
                st.exec = b.AND(src, old)
            elif op == "s_or_saveexec_b64":
                st.exec = b.OR(src, old)
            elif op == "s_andn2_saveexec_b64":
                st.exec = b.AND(src, b.NOT(old))
            elif op == "s_xor_saveexec_b64":
                st.exec = b.XOR(src, old)
            else:
                st.exec = self.atom(("unknown", site, "exec"))
            return
        if op.startswith("s_") and op.endswith("_b64") and ops:
            dst = ops[0]
            is_mask_dst = dst in ("exec", "vcc") or SPAIR.match(dst)
            binary = {"s_and_b64": b.AND, "s_or_b64": b.OR, "s_xor_b64": b.XOR,
                      "s_andn2_b64": lambda x, y: b.AND(x, b.NOT(y)),
                      "s_orn2_b64": lambda x, y: b.OR(x, b.NOT(y)),
                      "s_nand_b64": lambda x, y: b.NOT(b.AND(x, y)),
                      "s_nor_b64": lambda x, y: b.NOT(b.OR(x, y)),
                      "s_xnor_b64": lambda x, y: b.NOT(b.XOR(x, y))}
            if is_mask_dst and op in binary and len(ops) >= 3:
                self.set_mask(st, dst, binary[op](self.mask_value(st, ops[1], site),
                                                  self.mask_value(st, ops[2], site)))
                return
            if is_mask_dst and op == "s_mov_b64" and len(ops) >= 2:
                self.set_mask(st, dst, self.mask_value(st, ops[1], site))
                return
            if is_mask_dst and op == "s_not_b64" and len(ops) >= 2:
                self.set_mask(st, dst, b.NOT(self.mask_value(st, ops[1], site)))
                return
            if is_mask_dst and op == "s_cselect_b64" and len(ops) >= 3:
                u = self.atom(("uniform", site))
                self.set_mask(st, dst, b.OR(b.AND(u, self.mask_value(st, ops[1], site)),
                                            b.AND(b.NOT(u), self.mask_value(st, ops[2], site))))
                return
            if is_mask_dst:
                self.set_mask(st, dst, self.atom(("unknown", site, dst)))
                return
        if op.startswith("s_"):
            # other scalar writes: forget masks they overwrite
            if ops and not op.startswith(("s_cbranch", "s_branch", "s_waitcnt", "s_nop", "s_barrier", "s_cmp",
                                          "s_bitcmp", "s_setprio", "s_sleep", "s_endpgm", "s_sendmsg",
                                          "s_trap", "s_setpc", "s_getpc")):
                if ops[0] == "exec":
                    st.exec = self.atom(("unknown", site, "exec"))
                elif ops[0] == "vcc":
                    st.masks.pop("vcc", None)
                else:
                    self.kill_sregs(st, sregs(ops[0]))
            return
        if MEM.match(op):
            if check:
                for v in address_operands(op, ops, mods):
                    written = b.OR(st.defs.get(v, 0), st.defs.get("*", 0))
                    undefined = b.AND(st.exec, b.NOT(written))
                    if undefined != 0:
                        self.findings.append((op + " " + " ".join(ops), v))
            for v in access_dests(op, ops, mods):
                st.defs[v] = b.OR(st.defs.get(v, 0), st.exec)
            return
        if op.startswith("v_cmpx"):
            a = self.atom(("cmp", site))
            st.exec = b.AND(st.exec, a)
            if ops and (ops[0] == "vcc" or SPAIR.match(ops[0])):
                self.set_mask(st, ops[0], st.exec)
            return
        if op.startswith("v_cmp") or op.startswith("v_cmp_class"):
            a = self.atom(("cmp", site))
            dst = ops[0] if ops and (ops[0] == "vcc" or SPAIR.match(ops[0])) else "vcc"
            self.set_mask(st, dst, b.AND(st.exec, a))
            return
        if op.startswith("v_readlane") or op.startswith("v_readfirstlane"):
            if ops:
                self.kill_sregs(st, sregs(ops[0]))
            return
        if op.startswith("v_writelane"):
            return
        if op.startswith("v_") or op.startswith("ds_read") or op.startswith("ds_load") or op.startswith("ds_bpermute") \
                or op.startswith("ds_permute") or op.startswith("ds_swizzle") or op.startswith("ds_append") \
                or op.startswith("ds_consume"):
            if not ops:
                return
            for v in vregs(ops[0]):
                st.defs[v] = b.OR(st.defs.get(v, 0), st.exec)
            # VOP3 carry-out / second scalar destination (v_add_co_u32 v, s[..], ...)
            if len(ops) > 1 and (ops[1] == "vcc" or SPAIR.match(ops[1])) and op.startswith(
                    ("v_add_co", "v_sub_co", "v_subrev_co", "v_addc_co", "v_subb_co", "v_subbrev_co",
                     "v_mad_u64", "v_mad_i64", "v_div_scale")):
                self.set_mask(st, ops[1], b.AND(st.exec, self.atom(("carry", site))))
            return

    def edges(self, start, end, st):
        """Successor (index, state) pairs of the block [start, end) with exit state st."""
        b = self.bdd
        last = self.ins[end - 1]
        lab, op, payload = last
        out = []
        fall = end if end < len(self.ins) else None
        if op is None or not (op.startswith("s_branch") or op.startswith("s_cbranch") or op == "s_endpgm"):
            if fall is not None:
                out.append((fall, st))
            return out
        if op == "s_endpgm":
            return out
        target = self.labels.get(payload[0][0]) if payload[0] else None
        if op.startswith("s_branch"):
            if target is not None:
                out.append((target, st))
            return out

        def vacuous(s):
            e = s.copy()
            for v in list(e.defs):
                e.defs[v] = b.OR(e.defs[v], s.exec)
            e.defs["*"] = b.OR(e.defs.get("*", 0), s.exec)
            e.exec = 0
            return e

        if op == "s_cbranch_execz":
            if target is not None:
                out.append((target, vacuous(st)))
            if fall is not None:
                out.append((fall, st))
        elif op == "s_cbranch_execnz":
            if target is not None:
                out.append((target, st))
            if fall is not None:
                out.append((fall, vacuous(st)))
        elif op in ("s_cbranch_vccnz", "s_cbranch_vccz") and st.masks.get("vcc") == 0:
            # VCC is empty on this path: vccnz falls through, vccz jumps
            nxt = fall if op == "s_cbranch_vccnz" else target
            if nxt is not None:
                out.append((nxt, st))
        else:
            if target is not None:
                out.append((target, st))
            if fall is not None:
                out.append((fall, st))
        return out

    def merge(self, a, b_, site, widen=False):
        bdd = self.bdd
        m = State(bdd.OR(a.exec, b_.exec), {}, {})
        for k in set(a.masks) | set(b_.masks):
            va, vb = a.masks.get(k), b_.masks.get(k)
            if (site, k) in self.widened:
                m.masks[k] = self.atom(("widen", site, k))
            elif va is None or vb is None:
                m.masks[k] = self.atom(("merge", site, k))
            else:
                m.masks[k] = bdd.OR(va, vb)
                if widen and m.masks[k] != va:
                    self.widened.add((site, k))
                    m.masks[k] = self.atom(("widen", site, k))
        keys = set(a.defs) | set(b_.defs)
        sa, sb = a.defs.get("*", 0), b_.defs.get("*", 0)
        for k in keys:
            m.defs[k] = bdd.AND(bdd.OR(a.defs.get(k, 0), sa), bdd.OR(b_.defs.get(k, 0), sb))
        return m

    def run(self):
        """Worklist fixpoint; a block whose entry state is seen more than 40 times turns its
        still-changing masks into fixed atoms."""
        entry = State(1, {}, {0: 1})  # v0 holds the packed work-item ids in every lane
        states = {0: entry}
        visits = collections.Counter()
        self.widened = set()
        work = [0]
        total = 0
        while work:
            s = work.pop()
            total += 1
            if total > 200000:
                raise RuntimeError("%s: no fixpoint" % self.name)
            st = states[s].copy()
            end = self.block_of[s]
            for i in range(s, end):
                self.step(st, i, False)
            for t, ns in self.edges(s, end, st):
                old = states.get(t)
                visits[t] += 1
                new = self.merge(old, ns, t, visits[t] > 40) if old is not None else ns.copy()
                if old is None or new.key() != old.key():
                    states[t] = new
                    if t not in work:
                        work.append(t)
        for s in self.starts:
            if s in states:
                st = states[s].copy()
                for i in range(s, self.block_of[s]):
                    self.step(st, i, True)
        return self.findings


def check_assembly(text, skip=()):
    """{kernel: [(instruction, vgpr)]} of memory accesses whose address VGPR may be
    unwritten in some active lane; kernels whose name contains one of `skip` are left out."""
    out = {}
    for name, lines in kernels(text):
        if name.startswith("__") or name.startswith(".L") or any(k in name for k in skip):
            continue
        f = Kernel(name, lines, BDD()).run()
        if f:
            out[name] = f
    return out
