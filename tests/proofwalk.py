"""Field map of a serialized DeepFriProof<SexticExt> (test helper).

Walks the ark-serialize compressed bytes (crates/deep_ali/src/fri.rs:1235-1300;
usize as u64 LE, Vec as u64 length + items, Option as a 0/1 byte, field values
as 8 LE bytes) and returns (start, end, name) spans, so tests can tamper with
one named field at a time. Names are dotted paths with list positions, e.g.
"sp.3.f0p.path.1.2" (query 3's f0 opening, level 1, sibling 2).
"""
import struct


class Walker:
    def __init__(self, b):
        self.b = b
        self.pos = 0
        self.spans = []

    def take(self, n, name):
        if self.pos + n > len(self.b):
            raise ValueError("truncated at %s" % name)
        self.spans.append((self.pos, self.pos + n, name))
        v = self.b[self.pos:self.pos + n]
        self.pos += n
        return v

    def u64(self, name):
        return struct.unpack("<Q", self.take(8, name))[0]

    def flag(self, name):
        v = self.take(1, name)[0]
        if v > 1:
            raise ValueError("bad option flag at %s" % name)
        return v

    def opening(self, name):
        self.take(32, name + ".leaf")
        nl = self.u64(name + ".nlev")
        for l in range(nl):
            c = self.u64("%s.path.%d.len" % (name, l))
            for j in range(c):
                self.take(32, "%s.path.%d.%d" % (name, l, j))
        self.u64(name + ".index")


def walk(b):
    w = Walker(b)
    w.take(32, "root_f0")
    for i in range(w.u64("roots.len")):
        w.take(32, "roots.%d" % i)
    for l in range(w.u64("lp.len")):
        for q in range(w.u64("lp.%d.len" % l)):
            w.opening("lp.%d.%d" % (l, q))
    for q in range(w.u64("f0o.len")):
        w.opening("f0o.%d" % q)
    for q in range(w.u64("q.len")):
        for l in range(w.u64("q.%d.refs.len" % q)):
            for f in ("i", "child_pos", "parent_index", "parent_pos"):
                w.u64("q.%d.refs.%d.%s" % (q, l, f))
        for l in range(w.u64("q.%d.pay.len" % q)):
            for f in ("f", "s", "q"):
                w.take(48, "q.%d.pay.%d.%s" % (q, l, f))
        w.opening("q.%d.f0" % q)
        w.u64("q.%d.final_index" % q)
    for i in range(w.u64("fz.len")):
        w.take(48, "fz.%d" % i)
    for i in range(w.u64("fin.len")):
        w.take(48, "fin.%d" % i)
    w.u64("n0")
    w.u64("omega0")
    if w.flag("ct.some"):
        for t in range(w.u64("ct.len")):
            for k in range(w.u64("ct.%d.len" % t)):
                w.take(48, "ct.%d.%d" % (t, k))
    if w.flag("croot.some"):
        w.take(32, "croot")
    if w.flag("ce.some"):
        for l in range(w.u64("ce.len")):
            for j in range(w.u64("ce.%d.len" % l)):
                w.take(48, "ce.%d.%d" % (l, j))
    if w.flag("sp.some"):
        for q in range(w.u64("sp.len")):
            w.u64("sp.%d.base" % q)
            w.u64("sp.%d.raw" % q)
            for j in range(w.u64("sp.%d.fidx.len" % q)):
                w.u64("sp.%d.fidx.%d" % (q, j))
            for j in range(w.u64("sp.%d.fval.len" % q)):
                w.take(8, "sp.%d.fval.%d" % (q, j))
            w.opening("sp.%d.f0p" % q)
            w.take(48, "sp.%d.fnext" % q)
            if w.flag("sp.%d.l1.some" % q):
                w.opening("sp.%d.l1" % q)
    return w.spans, w.pos
