"""ctypes wrapper over the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg. The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U64P = C.POINTER(C.c_uint64)
U16P = C.POINTER(C.c_uint16)
U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_load.argtypes = [C.c_void_p, U8P, U64P, C.c_uint64]
        L.ref_set_names.argtypes = [C.c_void_p, C.c_char_p, U64P, C.c_uint64]
        L.ref_set_files.argtypes = [C.c_void_p, U32P, C.c_uint64]
        L.ref_p1.argtypes = [C.c_void_p, U64P]
        L.ref_p2.argtypes = [C.c_void_p, U16P, U32P]
        L.ref_bin_bytes.restype = C.c_uint64
        L.ref_bin_bytes.argtypes = [C.c_void_p, C.c_uint32, U8P]
        L.ref_bin_records.restype = C.c_uint64
        L.ref_bin_records.argtypes = [C.c_void_p, C.c_uint32, U64P, U64P, U64P]
        L.ref_p3.argtypes = [C.c_void_p, U64P]
        L.ref_keys.restype = C.c_uint64
        L.ref_keys.argtypes = [C.c_void_p, U64P, U64P]
        L.ref_per_bin_distinct.restype = C.c_uint64
        L.ref_per_bin_distinct.argtypes = [C.c_void_p, U64P]
        L.ref_mphf_export.restype = C.c_uint64
        L.ref_mphf_export.argtypes = [C.c_void_p, U64P]
        L.ref_mphf_lookup.argtypes = [C.c_void_p, U64P, U64P, C.c_uint64, U64P]
        L.ref_p4.argtypes = [C.c_void_p]
        L.ref_states.restype = C.c_uint64
        L.ref_states.argtypes = [C.c_void_p, U64P]
        L.ref_force_cas_failures.argtypes = [C.c_void_p, C.c_int64]
        L.ref_forced_cas_failures_left.restype = C.c_int64
        L.ref_forced_cas_failures_left.argtypes = [C.c_void_p]
        L.ref_p5.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int]
        L.ref_meta.argtypes = [C.c_void_p, U64P]
        L.ref_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ref_xxh3_u64.restype = C.c_uint64
        L.ref_xxh3_u64.argtypes = [C.c_uint64]
        L.ref_bucket_hash.restype = C.c_uint32
        L.ref_bucket_hash.argtypes = [C.c_uint64]
        L.ref_mmer_order.restype = C.c_uint32
        L.ref_mmer_order.argtypes = [C.c_uint32]
        L.ref_minimizer_len.argtypes = [C.c_int]
        L.ref_partition.argtypes = [U64P, C.c_uint32, U16P, U32P]
        L.ref_kmer_ops.argtypes = [C.c_int, C.c_char_p, C.c_int, U64P, C.c_char_p]
        L.ref_kmer_from_packed.argtypes = [C.c_int, U8P, C.c_uint64, U64P]
        L.ref_state_from_vertex.argtypes = [C.c_int] * 5
        L.ref_state_tables.argtypes = [U8P] * 5
        L.ref_sv_new.restype = C.c_void_p
        L.ref_sv_new.argtypes = [C.c_uint64]
        L.ref_sv_free.argtypes = [C.c_void_p]
        L.ref_sv_get.argtypes = [C.c_void_p, C.c_uint64]
        L.ref_sv_cas.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int]
        L.ref_sv_word.restype = C.c_uint64
        L.ref_sv_word.argtypes = [C.c_void_p, C.c_uint64]
        L.ref_set_target_seg.argtypes = [C.c_void_p, C.c_uint64]
        L.ref_dfa_step.argtypes = [C.c_int] * 5
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def pack_records(seqs):
    """list[bytes] -> (ascii uint8 array, offsets uint64 array of n+1)."""
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    if seqs:
        off[1:] = np.cumsum([len(s) for s in seqs], dtype=np.uint64)
    ascii_ = np.frombuffer(b"".join(seqs), dtype=np.uint8).copy() if seqs else np.zeros(0, np.uint8)
    if ascii_.size == 0:
        ascii_ = np.zeros(1, np.uint8)
    return ascii_, off


class RefPipeline:
    """The oracle pipeline for one input and one k (P1..P5)."""

    def __init__(self, k, num_bins=128, threads=1, target_seg=0):
        self.L = lib()
        self.k = k
        self.h = self.L.ref_create(k, num_bins, threads)
        if not self.h:
            raise ValueError("invalid k / num_bins")
        if target_seg:
            self.L.ref_set_target_seg(self.h, target_seg)

    def __del__(self):
        try:
            if self.h:
                self.L.ref_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def load(self, seqs, names=None):
        a, off = pack_records(seqs)
        self.load_arrays(a, off, names)

    def load_arrays(self, ascii_, off, names=None):
        """Records as one byte array and n+1 offsets (ascii_[off[r]:off[r+1]])."""
        a = np.ascontiguousarray(ascii_, np.uint8)
        if a.size == 0:
            a = np.zeros(1, np.uint8)
        off = np.ascontiguousarray(off, np.uint64)
        self._ascii, self._off = a, off
        self.L.ref_load(self.h, _p(a, U8P), _p(off, U64P), len(off) - 1)
        if names is not None:
            blob = "".join(names).encode()
            noff = np.zeros(len(names) + 1, np.uint64)
            noff[1:] = np.cumsum([len(n.encode()) for n in names])
            self.L.ref_set_names(self.h, blob, _p(noff, U64P), len(names))

    def set_files(self, file_idx):
        """Input file index of each record (cf_seq's Reference:<file+1>)."""
        f = np.ascontiguousarray(file_idx, np.uint32)
        if self.L.ref_set_files(self.h, _p(f, U32P), len(f)) != 0:
            raise ValueError("one file index per record")

    def p1(self):
        h = np.zeros(65536, np.uint64)
        self.L.ref_p1(self.h, _p(h, U64P))
        return h

    def p2(self):
        b2b = np.zeros(65536, np.uint16)
        nb = C.c_uint32(0)
        self.L.ref_p2(self.h, _p(b2b, U16P), C.byref(nb))
        self.num_bins = nb.value
        return b2b, nb.value

    def bin_bytes(self, b):
        n = self.L.ref_bin_bytes(self.h, b, None)
        out = np.zeros(max(n, 1), np.uint8)
        self.L.ref_bin_bytes(self.h, b, _p(out, U8P))
        return out[:n].tobytes()

    def bin_records(self, rw):
        """Every routed record as (bin, base count, rw words of bases MSB-first)."""
        n = self.L.ref_bin_records(self.h, rw, None, None, None)
        b = np.zeros(max(n, 1), np.uint64)
        ln = np.zeros(max(n, 1), np.uint64)
        w = np.zeros((max(n, 1), rw), np.uint64)
        self.L.ref_bin_records(self.h, rw, _p(b, U64P), _p(ln, U64P), _p(w, U64P))
        return b[:n], ln[:n], w[:n]

    def p3(self):
        d = C.c_uint64(0)
        self.L.ref_p3(self.h, C.byref(d))
        return d.value

    def keys(self):
        n = self.L.ref_keys(self.h, None, None)
        lo = np.zeros(max(n, 1), np.uint64)
        hi = np.zeros(max(n, 1), np.uint64)
        self.L.ref_keys(self.h, _p(lo, U64P), _p(hi, U64P))
        return lo[:n], hi[:n]

    def per_bin_distinct(self):
        n = self.L.ref_per_bin_distinct(self.h, None)
        out = np.zeros(max(n, 1), np.uint64)
        self.L.ref_per_bin_distinct(self.h, _p(out, U64P))
        return out[:n]

    def mphf_export(self):
        n = self.L.ref_mphf_export(self.h, None)
        out = np.zeros(n, np.uint64)
        self.L.ref_mphf_export(self.h, _p(out, U64P))
        return out

    def mphf_lookup(self, lo, hi=None):
        lo = np.ascontiguousarray(lo, np.uint64)
        out = np.zeros(max(len(lo), 1), np.uint64)
        hip = _p(np.ascontiguousarray(hi, np.uint64), U64P) if hi is not None else None
        self.L.ref_mphf_lookup(self.h, _p(lo, U64P), hip, len(lo), _p(out, U64P))
        return out[: len(lo)]

    def p4(self):
        self.L.ref_p4(self.h)

    def states(self):
        n = self.L.ref_states(self.h, None)
        out = np.zeros(max(n, 1), np.uint64)
        self.L.ref_states(self.h, _p(out, U64P))
        return out[:n]

    def force_cas_failures(self, n):
        self.L.ref_force_cas_failures(self.h, n)

    def forced_cas_failures_left(self):
        return self.L.ref_forced_cas_failures_left(self.h)

    def p5(self, prefix, input_desc="", poly_n_stretch=True, track_short=True):
        r = self.L.ref_p5(self.h, prefix.encode(), input_desc.encode(), int(poly_n_stretch), int(track_short))
        if r != 0:
            raise RuntimeError("oracle P5 failed to write output")

    def meta(self):
        m = np.zeros(5, np.uint64)
        self.L.ref_meta(self.h, _p(m, U64P))
        return dict(unitigs=int(m[0]), vertices=int(m[1]), max_len=int(m[2]), min_len=int(m[3]), sum_len=int(m[4]))

    def times(self):
        t = (C.c_double * 6)()
        self.L.ref_times(self.h, t)
        return dict(p1=t[0], p2=t[1], p3a=t[2], p3b=t[3], p4=t[4], p5=t[5])

    def run_p1_p4(self):
        self.p1()
        self.p2()
        d = self.p3()
        self.p4()
        return d


def state_of(words, idx):
    """5-bit state idx from 12-per-word packed words (state_vector.rs:36-42)."""
    idx = np.asarray(idx, np.uint64)
    w = words[(idx // 12).astype(np.int64)]
    return ((w >> ((idx % 12) * 5)) & 31).astype(np.uint8)
