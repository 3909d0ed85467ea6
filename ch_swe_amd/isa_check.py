"""Build-time check of the gfx950 assembly of libchswe's kernels.

For each kernel of the hipcc -S output: build the basic-block CFG, compute the
SGPRs that hold a defined value on every path to each instruction (the
compiler's "implicit-def" markers count as undefining a register), and report
address computations and memory instructions that read an SGPR that is not
defined on all paths.  Such a read means a load or store can use a stale base
address on some path.  This has happened with this toolchain: a stage kernel
that selected its store target from a runtime mode switch stored the last
RK4 stage through an SGPR pair that was only written on the other modes'
paths, and faulted on the GPU.  The build refuses a library with any finding.

Paths that cannot run are left out where the assembly shows it plainly: a
branch on an exec mask that was just set to zero (prune_zero_exec_edges).
"""
import re
import sys

NO_DEST = re.compile(r'^(global_store|buffer_store|ds_write|ds_store|flat_store|scratch_store|s_cmp|s_bitcmp|'
                     r's_cbranch|s_branch|s_waitcnt|s_barrier|s_endpgm|s_nop|s_setprio|s_sleep|buffer_wbl2|'
                     r'buffer_inv|s_dcache|s_icache|s_sendmsg|s_trap|s_setpc|s_wait|v_cmpx|s_set_gpr_idx|'
                     r'global_atomic|buffer_atomic|flat_atomic|ds_add|ds_max|ds_min|s_setreg|s_denorm|s_round)')
VOP3B = re.compile(r'^(v_add_co|v_sub_co|v_subrev_co|v_addc_co|v_subb_co|v_subbrev_co|v_mad_u64_u32|'
                   r'v_mad_i64_i32|v_div_scale)')
SREG = re.compile(r'\bs\[(\d+):(\d+)\]|\bs(\d+)\b')


def regs(op):
    out = []
    for m in SREG.finditer(op):
        if m.group(3) is not None:
            out.append(int(m.group(3)))
        else:
            out.extend(range(int(m.group(1)), int(m.group(2)) + 1))
    return out


def split_ops(rest):
    rest = rest.split(';')[0].strip()
    if not rest:
        return []
    parts, depth, cur = [], 0, ''
    for ch in rest:
        if ch == '[':
            depth += 1
        elif ch == ']':
            depth -= 1
        if ch == ',' and depth == 0:
            parts.append(cur.strip())
            cur = ''
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def _kills(mn, ops):
    if mn != '__kill__':
        return set()
    return set(int(x) for x in re.findall(r'sgpr(\d+)', ops[0]))


def _defs(mn, ops):
    if mn == '__kill__' or not ops or NO_DEST.match(mn):
        return set()
    d = set(regs(ops[0]))
    if VOP3B.match(mn) and len(ops) > 1:
        d.update(regs(ops[1]))
    return d


def _zero_step(z, mn, ops):
    """Known-zero SGPRs after one instruction."""
    z = z - _kills(mn, ops)
    d = _defs(mn, ops)
    z = z - d
    if mn in ('s_mov_b32', 's_mov_b64') and len(ops) == 2 and ops[1] == '0':
        z = z | d
    return z


def prune_zero_exec_edges(blocks, labels, preds):
    """Drop edges that cannot run.  A block whose last `s_and_saveexec_b64 d,
    m` is followed by nothing that writes exec up to its final
    `s_cbranch_execz L` / `s_cbranch_execnz L`, entered from a predecessor
    after which m is known to be zero, runs its branch with exec zero: execz
    always goes to L, execnz never does.  For each such predecessor the block
    is cloned with that one successor.  Known-zero SGPRs: a must-analysis over
    the CFG (`s_mov_b32/b64 s, 0` makes s zero, any other write or
    implicit-def unmakes it)."""
    ALL = set(range(0, 106))
    zout = [set(ALL) for _ in blocks]
    changed = True
    while changed:
        changed = False
        for i, b in enumerate(blocks):
            ps = [zout[p] for p in preds[i]]
            z = set() if i == 0 else (set.intersection(*ps) if ps else set(ALL))
            for ln, mn, ops in b['ins']:
                z = _zero_step(z, mn, ops)
            if z != zout[i]:
                zout[i] = z
                changed = True

    def writes_exec(mn, ops):
        return mn.startswith('v_cmpx') or 'saveexec' in mn or (ops and ops[0] in ('exec', 'exec_lo', 'exec_hi'))

    n0 = len(blocks)
    for i in range(n0):
        ins = blocks[i]['ins']
        if not ins or ins[-1][1] not in ('s_cbranch_execz', 's_cbranch_execnz'):
            continue
        k = len(ins) - 2
        while k >= 0 and not writes_exec(ins[k][1], ins[k][2]):
            k -= 1
        if k < 0 or ins[k][1] != 's_and_saveexec_b64' or len(ins[k][2]) != 2:
            continue
        mask = set(regs(ins[k][2][1]))
        if not mask:
            continue
        taken = labels[ins[-1][2][0]]
        fall = i + 1 if i + 1 < n0 else None
        only = taken if ins[-1][1] == 's_cbranch_execz' else fall
        if only is None:
            continue
        for p in list(preds[i]):
            if p == i:
                continue
            z = set(zout[p])
            for ln, mn, ops in ins[:k]:
                z = _zero_step(z, mn, ops)
            if mask <= z:
                clone = {'label': blocks[i]['label'] + f'/exec0-from-{p}', 'ins': list(ins), 'succ': [only]}
                blocks.append(clone)
                blocks[p]['succ'] = [len(blocks) - 1 if s == i else s for s in blocks[p]['succ']]


def analyze(name, body, n_entry):
    # instructions: (lineno, mnemonic, ops)
    blocks, labels, cur = [], {}, None
    for ln, text in body:
        t = text.strip()
        if not t or t.startswith(';') or t.startswith('.') and not re.match(r'^\.LBB\S+:', t):
            if t.startswith('; %bb.'):
                cur = {'label': t, 'ins': [], 'succ': []}
                blocks.append(cur)
            m = re.match(r'^;\s*implicit-def:\s*(.*)$', t)
            if m and cur is not None:
                for tok in m.group(1).split(','):
                    tok = tok.strip()
                    if tok.startswith('$sgpr'):
                        cur['ins'].append((ln, '__kill__', [tok]))
            continue
        m = re.match(r'^(\.LBB\S+):', t)
        if m:
            cur = {'label': m.group(1), 'ins': [], 'succ': []}
            labels[m.group(1)] = len(blocks)
            blocks.append(cur)
            continue
        if cur is None:
            cur = {'label': 'entry', 'ins': [], 'succ': []}
            blocks.append(cur)
        parts = t.split(None, 1)
        mn = parts[0]
        ops = split_ops(parts[1]) if len(parts) > 1 else []
        cur['ins'].append((ln, mn, ops))
        if mn.startswith('s_cbranch') or mn == 's_branch' or mn == 's_endpgm' or mn.startswith('s_setpc'):
            cur = {'label': f'after{ln}', 'ins': [], 'succ': []}
            blocks.append(cur)
    # successors
    for i, b in enumerate(blocks):
        last = b['ins'][-1] if b['ins'] else None
        if last and last[1] == 's_branch':
            b['succ'] = [labels[last[2][0]]]
        elif last and last[1].startswith('s_cbranch'):
            b['succ'] = [labels[last[2][0]]] + ([i + 1] if i + 1 < len(blocks) else [])
        elif last and (last[1] == 's_endpgm' or last[1].startswith('s_setpc')):
            b['succ'] = []
        else:
            b['succ'] = [i + 1] if i + 1 < len(blocks) else []
    preds = [[] for _ in blocks]
    for i, b in enumerate(blocks):
        for s in b['succ']:
            preds[s].append(i)
    prune_zero_exec_edges(blocks, labels, preds)
    preds = [[] for _ in blocks]
    for i, b in enumerate(blocks):
        for s in b['succ']:
            preds[s].append(i)

    def kills_of(mn, ops):
        if mn != '__kill__':
            return set()
        return set(int(x) for x in re.findall(r'sgpr(\d+)', ops[0]))

    def defs_of(mn, ops):
        d = set()
        if mn == '__kill__':
            return d
        if not ops or NO_DEST.match(mn):
            if mn.startswith('global_atomic') or mn.startswith('buffer_atomic') or mn.startswith('flat_atomic'):
                pass  # returning atomics write a VGPR only
            return d
        d.update(regs(ops[0]))
        if VOP3B.match(mn) and len(ops) > 1:
            d.update(regs(ops[1]))
        return d

    ALL = set(range(0, 106))
    out = [set(ALL) for _ in blocks]
    entry_set = set(range(0, n_entry))
    changed = True
    while changed:
        changed = False
        for i, b in enumerate(blocks):
            if i == 0:
                s = set(entry_set)
            else:
                ps = [out[p] for p in preds[i]]
                s = set.intersection(*ps) if ps else set(ALL)  # unreachable: top
            for ln, mn, ops in b['ins']:
                s -= kills_of(mn, ops)
                s |= defs_of(mn, ops)
            if s != out[i]:
                out[i] = s
                changed = True
    reports = []
    for i, b in enumerate(blocks):
        if i == 0:
            s = set(entry_set)
        else:
            ps = [out[p] for p in preds[i]]
            s = set.intersection(*ps) if ps else set(ALL)
        for ln, mn, ops in b['ins']:
            if mn == '__kill__':
                s -= kills_of(mn, ops)
                continue
            d = defs_of(mn, ops)
            srcs = ops[1:] if d or (ops and not NO_DEST.match(mn)) else ops
            if NO_DEST.match(mn):
                srcs = ops
            if VOP3B.match(mn):
                srcs = ops[2:]
            if mn.startswith('s_cbranch') or mn == 's_branch':
                srcs = []
            addr = re.match(r'^(global_|s_load|s_buffer_load|buffer_|flat_|scratch_|v_lshl_add_u64|'
                            r'v_add_co_u32|v_addc_co_u32|s_add_u32|s_addc_u32)', mn)
            for op in srcs if addr else []:
                for r in regs(op):
                    if r not in s:
                        reports.append((ln, mn, ops, r, b['label']))
            s |= d
    return reports


def check_file(path, pat=r'.'):
    """{kernel: [(line, mnemonic, operands, sgpr, block)]} for kernels matching
    pat.  An out-of-line device function (not a kernel) is reported as one
    finding: the analysis assumes a kernel's entry state, and the fused
    kernels are meant to inline everything they call."""
    lines = open(path).read().split('\n')
    kernels = set(re.findall(r'^\s*\.amdhsa_kernel\s+(\S+)', '\n'.join(lines), re.M))
    kern = []
    for i, l in enumerate(lines):
        m = re.match(r'^(_Z\S+|k_\S+):\s*;\s*@', l)
        if m:
            kern.append((m.group(1), i))
    found = {}
    for name, start in kern:
        if not re.search(pat, name):
            continue
        if kernels and name not in kernels:
            found[name] = [(start + 1, 'out-of-line device function', [], -1, 'entry')]
            continue
        end = start
        while end < len(lines) and not lines[end].startswith('.Lfunc_end'):
            end += 1
        body = [(j + 1, lines[j]) for j in range(start + 1, end)]
        rep = analyze(name, body, 20)
        if rep:
            found[name] = rep
    return found


def main():
    path = sys.argv[1]
    pat = sys.argv[2] if len(sys.argv) > 2 else r'k_stage|k_records'
    lines = open(path).read().split('\n')
    kern = []
    for i, l in enumerate(lines):
        m = re.match(r'^(_Z\S+):\s*;\s*@', l)
        if m:
            kern.append((m.group(1), i))
    for k, (name, start) in enumerate(kern):
        if not re.search(pat, name):
            continue
        end = start
        while end < len(lines) and not lines[end].startswith('.Lfunc_end'):
            end += 1
        body = [(j + 1, lines[j]) for j in range(start + 1, end)]
        # user + system SGPRs available at entry
        meta = '\n'.join(lines[end:end + 200])
        mu = re.search(r'\.amdhsa_user_sgpr_count (\d+)', meta)
        n_entry = (int(mu.group(1)) if mu else 16) + 4
        rep = analyze(name, body, n_entry)
        print(f'{name[:80]}: {len(rep)} reads of possibly-undefined SGPRs')
        for ln, mn, ops, r, lab in rep[:int(sys.argv[3]) if len(sys.argv) > 3 else 8]:
            print(f'    line {ln} [{lab}] {mn} {", ".join(ops)}   (s{r})')


if __name__ == '__main__':
    main()
