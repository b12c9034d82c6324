"""Python bindings (ctypes) over libcf1g.so, the MI355X phase library.

The compute path is the HIP library; this module only marshals buffers. It
never falls back to a CPU implementation: if libcf1g.so is missing or no
gfx950 device is visible, every entry point raises.

Method names follow the reference's phase functions (src/pipeline.rs:66-118):
count_minimizer_histogram (P1), partition_minimizers_and_route (P2),
build_mphf (P3), classify_vertices (P4).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcf1g.so")
_LIB = None

U64P = C.POINTER(C.c_uint64)
U32P = C.POINTER(C.c_uint32)
U16P = C.POINTER(C.c_uint16)
U8P = C.POINTER(C.c_uint8)

CF1G_OK = 0
ERRORS = {-1: "CF1G_E_INVALID", -2: "CF1G_E_ORDER", -3: "CF1G_E_HIP", -4: "CF1G_E_NODEVICE", -5: "CF1G_E_INTERNAL"}
NUM_BUCKETS = 65536

# Every symbol include/cf1g.h declares.
EXPORTS = [
    "cf1g_create", "cf1g_destroy", "cf1g_last_error", "cf1g_get_info", "cf1g_load_records",
    "cf1g_load_fasta", "cf1g_loaded_records", "cf1g_stream_export",
    "cf1g_bins_write", "cf1g_dedup_write", "cf1g_save", "cf1g_restore",
    "cf1g_p1_histogram", "cf1g_p2_partition_route", "cf1g_p3_build", "cf1g_p4_classify",
    "cf1g_p4_classify_cas", "cf1g_autotune", "cf1g_run_p1_p4", "cf1g_states_export", "cf1g_keys_export", "cf1g_mphf_export",
    "cf1g_mphf_lookup", "cf1g_mphf_lookup_device", "cf1g_annotate", "cf1g_unitig_spans", "cf1g_record_starts",
    "cf1g_set_record_files", "cf1g_record_files", "cf1g_mphf_lookup_noalloc", "cf1g_p1_stream",
    "cf1g_run_p1_p4_stream",
    "cf1g_bins_export", "cf1g_p5_write", "cf1g_sync", "cf1g_set_graphs", "cf1g_set_profiling", "cf1g_kernel_times",
    "cf1g_device_keys", "cf1g_lidx_train", "cf1g_lidx_lookup", "cf1g_lidx_lookup_device", "cf1g_lidx_export",
    "cf1g_lidx_gradients", "cf1g_classify_via", "cf1g_select_index", "cf1g_selected_index",
]

# Learned-index records (the layouts of include/cf1g.h).
ROOT_DTYPE = np.dtype([("A", "<f4"), ("B", "<f4"), ("leaf0", "<u4"), ("p", "<u4")])
LEAF_DTYPE = np.dtype([("S", "<f4"), ("I", "<f4")])


class Cf1gError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class CfParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("num_bins", C.c_uint32), ("m", C.c_uint32), ("p4_chunk", C.c_uint32),
                ("p4_wgs_per_cu", C.c_uint32), ("p3_target_seg", C.c_uint32), ("stream_chunk_log2", C.c_uint32),
                ("reserved", C.c_uint32)]


class CfInfo(C.Structure):
    _fields_ = [("n_records", C.c_uint64), ("stream_len", C.c_uint64), ("n_kmers", C.c_uint64),
                ("n_superkmers", C.c_uint64), ("n_distinct", C.c_uint64), ("n_state_words", C.c_uint64),
                ("mphf_words", C.c_uint64), ("num_bins", C.c_uint32), ("seg_bits", C.c_uint32),
                ("mphf_big_segments", C.c_uint32), ("n_oversized", C.c_uint32), ("k", C.c_uint32), ("m", C.c_uint32),
                ("graph_captures", C.c_uint64), ("graph_replays", C.c_uint64)]


class IngestStats(C.Structure):
    _fields_ = [("n_records", C.c_uint64), ("bases", C.c_uint64), ("stream_bytes", C.c_uint64),
                ("parse_ms", C.c_double), ("pack_ms", C.c_double), ("h2d_ms", C.c_double)]


class CasStats(C.Structure):
    _fields_ = [("occurrences", C.c_uint64), ("cas_issued", C.c_uint64), ("cas_failed", C.c_uint64),
                ("n_chunks", C.c_uint64), ("chunk", C.c_uint32), ("wgs_per_cu", C.c_uint32), ("grid", C.c_uint32),
                ("reserved", C.c_uint32), ("kernel_ms", C.c_double)]


CAS_FORCE_RETRY = 1


class TuneEntry(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("a", C.c_uint32), ("b", C.c_uint32), ("c", C.c_uint32), ("ms", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("num_bins", C.c_uint32), ("p3_target_seg", C.c_uint32), ("p4_chunk", C.c_uint32),
                ("p4_wgs_per_cu", C.c_uint32), ("step_ms", C.c_double), ("p4_cas_ms", C.c_double),
                ("n_entries", C.c_uint32), ("reserved", C.c_uint32)]


class LidxParams(C.Structure):
    _fields_ = [("leaf_bits", C.c_uint32), ("root_bits", C.c_uint32), ("epochs", C.c_uint32), ("lr", C.c_float),
                ("max_window", C.c_uint32), ("untrained", C.c_uint32), ("line_bytes", C.c_uint32), ("line_keys", C.c_uint32)]


class LidxInfo(C.Structure):
    _fields_ = [("n_keys", C.c_uint64), ("n_leaves", C.c_uint64), ("n_pieces", C.c_uint64),
                ("fallback_leaves", C.c_uint64), ("fallback_keys", C.c_uint64), ("n_lines", C.c_uint64),
                ("overflow_lines", C.c_uint64), ("search_keys", C.c_uint64), ("line_bytes", C.c_uint32),
                ("line_keys", C.c_uint32), ("max_err_lo", C.c_uint32), ("max_err_hi", C.c_uint32),
                ("mean_abs_err", C.c_double), ("line_bits_per_key", C.c_double), ("bytes_total", C.c_uint64),
                ("root_loss", C.c_double * 2), ("leaf_loss", C.c_double * 2), ("sort_ms", C.c_double),
                ("train_ms", C.c_double)]


class LidxShape(C.Structure):
    _fields_ = [("n_keys", C.c_uint64), ("n_pieces", C.c_uint32), ("n_leaves", C.c_uint32), ("root_shift", C.c_int32),
                ("line_bytes", C.c_uint32), ("line_keys", C.c_uint32), ("n_lines", C.c_uint32), ("key_bits", C.c_uint32),
                ("warp_groups", C.c_uint32)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C cf1-rs_amd`)")
        L = C.CDLL(LIB_PATH)
        L.cf1g_create.argtypes = [C.POINTER(CfParams), C.POINTER(C.c_void_p)]
        L.cf1g_destroy.argtypes = [C.c_void_p]
        L.cf1g_last_error.restype = C.c_char_p
        L.cf1g_last_error.argtypes = [C.c_void_p]
        L.cf1g_get_info.argtypes = [C.c_void_p, C.POINTER(CfInfo)]
        L.cf1g_load_records.argtypes = [C.c_void_p, U8P, U64P, C.c_uint64]
        L.cf1g_load_fasta.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.POINTER(IngestStats)]
        L.cf1g_loaded_records.argtypes = [C.c_void_p, C.POINTER(U8P), C.POINTER(U64P), U64P, C.POINTER(C.c_char_p),
                                          C.POINTER(U64P)]
        L.cf1g_stream_export.argtypes = [C.c_void_p, U64P, U64P, C.c_uint64]
        L.cf1g_bins_write.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32, U64P]
        L.cf1g_dedup_write.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32]
        L.cf1g_save.argtypes = [C.c_void_p, C.c_char_p]
        L.cf1g_restore.argtypes = [C.c_void_p, C.c_char_p]
        L.cf1g_p1_histogram.argtypes = [C.c_void_p, U64P]
        L.cf1g_p1_stream.argtypes = [C.c_void_p, U8P, U64P, C.c_uint64, C.c_uint32, U64P]
        L.cf1g_run_p1_p4_stream.argtypes = [C.c_void_p, U8P, U64P, C.c_uint64, C.c_uint32, U64P]
        L.cf1g_p2_partition_route.argtypes = [C.c_void_p, U64P, C.c_uint32, U16P, U32P]
        L.cf1g_p3_build.argtypes = [C.c_void_p, U64P]
        L.cf1g_p4_classify.argtypes = [C.c_void_p]
        L.cf1g_p4_classify_cas.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(CasStats)]
        L.cf1g_autotune.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(TuneEntry), C.c_uint32, C.POINTER(TuneResult)]
        L.cf1g_run_p1_p4.argtypes = [C.c_void_p, U64P]
        L.cf1g_states_export.argtypes = [C.c_void_p, U64P, C.c_uint64]
        L.cf1g_keys_export.argtypes = [C.c_void_p, U64P, C.c_uint64]
        L.cf1g_mphf_export.argtypes = [C.c_void_p, U64P, U64P]
        L.cf1g_mphf_lookup.argtypes = [C.c_void_p, U64P, C.c_uint64, U64P]
        L.cf1g_mphf_lookup_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.cf1g_mphf_lookup_noalloc.argtypes = [C.c_void_p, U64P, C.c_uint64, U64P]
        L.cf1g_annotate.argtypes = [C.c_void_p, U8P, C.c_uint64]
        L.cf1g_unitig_spans.argtypes = [C.c_void_p, U64P, U64P, U64P, U64P, U8P, C.c_uint64]
        L.cf1g_lidx_train.argtypes = [C.c_void_p, C.POINTER(LidxParams), C.POINTER(LidxInfo)]
        L.cf1g_lidx_lookup.argtypes = [C.c_void_p, U64P, C.c_uint64, U64P]
        L.cf1g_lidx_lookup_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.cf1g_lidx_export.argtypes = [C.c_void_p, C.POINTER(LidxShape), U64P, U32P, C.c_void_p, C.c_void_p, U32P, C.c_void_p]
        F32P = C.POINTER(C.c_float)
        F64P = C.POINTER(C.c_double)
        L.cf1g_lidx_gradients.argtypes = [C.c_void_p, C.c_int, F32P, F32P, C.c_uint64, F64P, F64P, F64P, U64P, F32P, F32P]
        L.cf1g_classify_via.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.cf1g_select_index.argtypes = [C.c_void_p, C.c_int]
        L.cf1g_selected_index.argtypes = [C.c_void_p]
        L.cf1g_record_starts.argtypes = [C.c_void_p, U64P, C.c_uint64]
        L.cf1g_set_record_files.argtypes = [C.c_void_p, U32P, C.c_uint64]
        L.cf1g_record_files.argtypes = [C.c_void_p, U32P, C.c_uint64]
        L.cf1g_bins_export.argtypes = [C.c_void_p, U32P, U64P, U64P, C.c_uint64]
        L.cf1g_p5_write.argtypes = [C.c_void_p, U8P, U64P, C.c_uint64, C.c_char_p, U64P, C.c_char_p, C.c_char_p,
                                    C.c_int, C.c_int, C.c_int, U64P]
        L.cf1g_sync.argtypes = [C.c_void_p]
        L.cf1g_set_graphs.argtypes = [C.c_void_p, C.c_int]
        L.cf1g_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.cf1g_kernel_times.restype = C.c_size_t
        L.cf1g_kernel_times.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.c_int]
        L.cf1g_device_keys.restype = C.c_void_p
        L.cf1g_device_keys.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def pack_records(seqs):
    """list[bytes] -> (ascii uint8, rec_off uint64[n+1])."""
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    if seqs:
        off[1:] = np.cumsum([len(s) for s in seqs], dtype=np.uint64)
    a = np.frombuffer(b"".join(seqs), dtype=np.uint8).copy() if seqs else np.zeros(0, np.uint8)
    if a.size == 0:
        a = np.zeros(1, np.uint8)
    return a, off


def minimizer_len(k):
    """src/params.rs:221-228."""
    m = min(k // 2, 15)
    if m % 2 == 0:
        m = max(m - 1, 1)
    return m


class Cf1g:
    """One device context: the P1..P4 phases on the loaded records."""

    def __init__(self, k=31, num_bins=128, m=0, p4_chunk=0, p4_wgs_per_cu=0, p3_target_seg=0, stream_chunk_log2=0):
        self.L = lib()
        self.k = k
        p = CfParams(k=k, num_bins=num_bins, m=m, p4_chunk=p4_chunk, p4_wgs_per_cu=p4_wgs_per_cu,
                     p3_target_seg=p3_target_seg, stream_chunk_log2=stream_chunk_log2)
        h = C.c_void_p()
        rc = self.L.cf1g_create(C.byref(p), C.byref(h))
        self.h = h
        self.num_bins_req = num_bins
        if rc != CF1G_OK:
            msg = self.L.cf1g_last_error(h).decode() if h else "cf1g_create failed"
            self.close()
            raise Cf1gError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.cf1g_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != CF1G_OK:
            raise Cf1gError(rc, self.L.cf1g_last_error(self.h).decode())

    # ---- input
    def load_records(self, seqs=None, ascii=None, rec_off=None):
        if seqs is not None:
            ascii, rec_off = pack_records(seqs)
        self._ascii = np.ascontiguousarray(ascii, np.uint8)
        self._off = np.ascontiguousarray(rec_off, np.uint64)
        self._from_fasta = False
        self._check(self.L.cf1g_load_records(self.h, _p(self._ascii, U8P), _p(self._off, U64P), len(self._off) - 1))

    # ---- phases (src/pipeline.rs:77-103)
    def load_fasta(self, paths, threads=8):
        """Parse FASTA files (plain or .gz) on the host, pack the 2-bit stream
        there and copy it to HBM (cf1g_load_fasta); returns ingest stats."""
        if isinstance(paths, (str, bytes, os.PathLike)):
            paths = [paths]
        arr = (C.c_char_p * max(len(paths), 1))(*[os.fsencode(p) for p in paths])
        o = IngestStats()
        self._from_fasta = False
        self._check(self.L.cf1g_load_fasta(self.h, arr, len(paths), threads, C.byref(o)))
        self._from_fasta = True
        return {f: getattr(o, f) for f, _ in IngestStats._fields_}

    def loaded_records(self):
        """(ascii, rec_off, names) of the last load_fasta (copies)."""
        a, ro, no = U8P(), U64P(), U64P()
        n = C.c_uint64(0)
        nm = C.c_char_p()
        self._check(self.L.cf1g_loaded_records(self.h, C.byref(a), C.byref(ro), C.byref(n), C.byref(nm), C.byref(no)))
        n = n.value
        rec_off = np.ctypeslib.as_array(ro, shape=(n + 1,)).copy()
        name_off = np.ctypeslib.as_array(no, shape=(n + 1,)).copy()
        ascii_ = np.ctypeslib.as_array(a, shape=(max(int(rec_off[-1]), 1),))[:int(rec_off[-1])].copy()
        blob = C.string_at(nm, int(name_off[-1])) if name_off[-1] else b""
        names = [blob[int(name_off[i]):int(name_off[i + 1])].decode() for i in range(n)]
        return ascii_, rec_off, names

    def stream_export(self):
        """(seq, nmask) words of the packed stream in HBM."""
        nw = (self.info()["stream_len"] + 63) // 64
        seq = np.zeros(2 * nw + 1, np.uint64)
        nmask = np.zeros(nw + 1, np.uint64)
        self._check(self.L.cf1g_stream_export(self.h, _p(seq, U64P), _p(nmask, U64P), nw))
        return seq[:2 * nw], nmask[:nw]

    # ---- the reference's intermediate files, and resume
    def bins_write(self, directory, threads=8):
        n = C.c_uint64(0)
        self._check(self.L.cf1g_bins_write(self.h, os.fsencode(directory), threads, C.byref(n)))
        return n.value

    def dedup_write(self, path, threads=8):
        self._check(self.L.cf1g_dedup_write(self.h, os.fsencode(path), threads))

    def save(self, path):
        self._check(self.L.cf1g_save(self.h, os.fsencode(path)))

    def restore(self, path):
        self._check(self.L.cf1g_restore(self.h, os.fsencode(path)))

    def p1_stream(self, seqs=None, ascii=None, rec_off=None, threads=0, want=True):
        """Load the records and run P1 while they stream to the device
        (cf1g_p1_stream): the same products as load_records + P1."""
        if seqs is not None:
            ascii, rec_off = pack_records(seqs)
        self._ascii = np.ascontiguousarray(ascii, np.uint8)
        self._off = np.ascontiguousarray(rec_off, np.uint64)
        self._from_fasta = False
        h = np.zeros(NUM_BUCKETS, np.uint64) if want else None
        self._check(self.L.cf1g_p1_stream(self.h, _p(self._ascii, U8P), _p(self._off, U64P), len(self._off) - 1, threads,
                                          _p(h, U64P) if want else None))
        return h

    def run_p1_p4_stream(self, ascii, rec_off, threads=0):
        """P1..P4 from records in host memory (cf1g_run_p1_p4_stream)."""
        self._ascii = np.ascontiguousarray(ascii, np.uint8)
        self._off = np.ascontiguousarray(rec_off, np.uint64)
        self._from_fasta = False
        d = C.c_uint64(0)
        self._check(self.L.cf1g_run_p1_p4_stream(self.h, _p(self._ascii, U8P), _p(self._off, U64P), len(self._off) - 1, threads,
                                                 C.byref(d)))
        return d.value

    def count_minimizer_histogram(self, want=True):
        h = np.zeros(NUM_BUCKETS, np.uint64) if want else None
        self._check(self.L.cf1g_p1_histogram(self.h, _p(h, U64P) if want else None))
        return h

    def partition_minimizers_and_route(self, hist=None, num_bins=None):
        b2b = np.zeros(NUM_BUCKETS, np.uint16)
        nb = C.c_uint32(0)
        hp = _p(np.ascontiguousarray(hist, np.uint64), U64P) if hist is not None else None
        self._check(self.L.cf1g_p2_partition_route(self.h, hp, num_bins or self.num_bins_req, _p(b2b, U16P), C.byref(nb)))
        return b2b, nb.value

    def build_mphf(self):
        d = C.c_uint64(0)
        self._check(self.L.cf1g_p3_build(self.h, C.byref(d)))
        return d.value

    def classify_vertices(self):
        self._check(self.L.cf1g_p4_classify(self.h))

    def classify_vertices_cas(self, force_retry=False):
        """P4 in the reference's form: per-occurrence MPHF lookup and CAS
        join into a zeroed state vector; returns the kernel's statistics."""
        o = CasStats()
        self._check(self.L.cf1g_p4_classify_cas(self.h, CAS_FORCE_RETRY if force_retry else 0, C.byref(o)))
        return {f: getattr(o, f) for f, _ in CasStats._fields_ if f != "reserved"}

    def autotune(self, reps=2):
        """Time the candidate launch configurations on the loaded records and
        keep the fastest (cf1g_autotune); returns the choice and every
        measurement."""
        ents = (TuneEntry * 64)()
        o = TuneResult()
        self._check(self.L.cf1g_autotune(self.h, reps, ents, 64, C.byref(o)))
        res = {f: getattr(o, f) for f, _ in TuneResult._fields_ if f != "reserved"}
        kinds = {0: ("step", "num_bins", "p3_target_seg"), 1: ("p4_cas", "p4_chunk", "p4_wgs_per_cu")}
        table = []
        for e in ents[:o.n_entries]:
            kind, an, bn = kinds[e.kind]
            table.append({"kind": kind, an: e.a, bn: e.b, "ms": round(e.ms, 4)})
        res["table"] = table
        return res

    def run_p1_p4(self):
        d = C.c_uint64(0)
        self._check(self.L.cf1g_run_p1_p4(self.h, C.byref(d)))
        return d.value

    # ---- products
    def info(self):
        i = CfInfo()
        self._check(self.L.cf1g_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in CfInfo._fields_}

    def states(self):
        n = self.info()["n_state_words"]
        out = np.zeros(max(n, 1), np.uint64)
        self._check(self.L.cf1g_states_export(self.h, _p(out, U64P), out.size))
        return out[:n]

    def keys(self):
        inf = self.info()
        wpk = 2 if self.k > 32 else 1
        out = np.zeros(max(inf["n_distinct"] * wpk, 1), np.uint64)
        self._check(self.L.cf1g_keys_export(self.h, _p(out, U64P), out.size))
        out = out[: inf["n_distinct"] * wpk]
        if wpk == 2:
            return out[0::2].copy(), out[1::2].copy()
        return out, np.zeros_like(out)

    def mphf_export(self):
        n = C.c_uint64(0)
        self._check(self.L.cf1g_mphf_export(self.h, None, C.byref(n)))
        out = np.zeros(n.value, np.uint64)
        self._check(self.L.cf1g_mphf_export(self.h, _p(out, U64P), C.byref(n)))
        return out

    def mphf_lookup(self, lo, hi=None):
        lo = np.ascontiguousarray(lo, np.uint64)
        if self.k > 32:
            keys = np.empty(2 * len(lo), np.uint64)
            keys[0::2] = lo
            keys[1::2] = np.ascontiguousarray(hi, np.uint64)
        else:
            keys = lo
        out = np.zeros(max(len(lo), 1), np.uint64)
        self._check(self.L.cf1g_mphf_lookup(self.h, _p(keys, U64P), len(lo), _p(out, U64P)))
        return out[: len(lo)]

    def annotate(self):
        n = self.info()["stream_len"]
        out = np.zeros(max(n, 1), np.uint8)
        self._check(self.L.cf1g_annotate(self.h, _p(out, U8P), out.size))
        return out[:n]

    def unitig_spans(self):
        """(start, end, id, flags) per maximal unitig in stream order
        (cf1g_unitig_spans)."""
        n = C.c_uint64(0)
        self._check(self.L.cf1g_unitig_spans(self.h, C.byref(n), None, None, None, None, 0))
        u = n.value
        s, e, i = (np.zeros(u + 1, np.uint64) for _ in range(3))
        f = np.zeros(u + 1, np.uint8)
        self._check(self.L.cf1g_unitig_spans(self.h, C.byref(n), _p(s, U64P), _p(e, U64P), _p(i, U64P), _p(f, U8P), u + 1))
        return s[:u], e[:u], i[:u], f[:u]

    def record_starts(self):
        n = self.info()["n_records"] + 1
        out = np.zeros(n, np.uint64)
        self._check(self.L.cf1g_record_starts(self.h, _p(out, U64P), n))
        return out

    def set_record_files(self, file_idx):
        """Input file (0-based) of each loaded record: cf_seq's Reference:<file+1>."""
        f = np.ascontiguousarray(file_idx, np.uint32)
        self._check(self.L.cf1g_set_record_files(self.h, _p(f, U32P), len(f)))

    def record_files(self):
        n = self.info()["n_records"]
        out = np.zeros(max(n, 1), np.uint32)
        self._check(self.L.cf1g_record_files(self.h, _p(out, U32P), max(n, 1)))
        return out[:n]

    def bins_export(self):
        inf = self.info()
        rw = C.c_uint32(0)
        bs = np.zeros(inf["num_bins"] + 1, np.uint64)
        self._check(self.L.cf1g_bins_export(self.h, C.byref(rw), _p(bs, U64P), None, 0))
        recs = np.zeros(max(inf["n_superkmers"] * rw.value, 1), np.uint64)
        self._check(self.L.cf1g_bins_export(self.h, None, None, _p(recs, U64P), recs.size))
        return rw.value, bs, recs[: inf["n_superkmers"] * rw.value].reshape(-1, rw.value)

    def p5_write(self, prefix, names=None, input_desc="", poly_n_stretch=True, track_short=True, threads=4):
        """Host P5: extract_and_output + write_json (src/output.rs:152-354, 785-834).
        After load_fasta the context's records (and their ids) are used."""
        if getattr(self, "_from_fasta", False):
            self._ascii, self._off, fnames = self.loaded_records()
            names = fnames if names is None else names
        if names is not None:
            enc = [n.encode() for n in names]
            blob = b"".join(enc)
            noff = np.zeros(len(enc) + 1, np.uint64)
            noff[1:] = np.cumsum([len(e) for e in enc])
            nptr = _p(noff, U64P)
        else:
            blob, nptr = None, None
        meta = np.zeros(5, np.uint64)
        self._check(self.L.cf1g_p5_write(self.h, _p(self._ascii, U8P), _p(self._off, U64P), len(self._off) - 1, blob, nptr,
                                         prefix.encode(), input_desc.encode(), int(poly_n_stretch), int(track_short),
                                         threads, _p(meta, U64P)))
        return dict(unitigs=int(meta[0]), vertices=int(meta[1]), max_len=int(meta[2]), min_len=int(meta[3]),
                    sum_len=int(meta[4]))

    # ---- measurement
    def sync(self):
        self._check(self.L.cf1g_sync(self.h))

    def set_graphs(self, on):
        self._check(self.L.cf1g_set_graphs(self.h, 1 if on else 0))

    def set_profiling(self, on):
        self._check(self.L.cf1g_set_profiling(self.h, int(on)))

    def kernel_times(self, reset=True):
        n = self.L.cf1g_kernel_times(self.h, None, 0, 0)
        buf = C.create_string_buffer(max(n, 1) + 256)
        self.L.cf1g_kernel_times(self.h, buf, len(buf), int(reset))
        res = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            res[name] = (int(cnt), float(ms))
        return res


def _lidx_methods():
    """Learned rank index methods of Cf1g (kept together here)."""

    def lidx_train(self, leaf_bits=0, root_bits=0, epochs=0, lr=0.0, max_window=0, untrained=False, line_bytes=0,
                   line_keys=0):
        p = LidxParams(leaf_bits, root_bits, epochs, lr, max_window, 1 if untrained else 0, line_bytes, line_keys)
        o = LidxInfo()
        self._check(self.L.cf1g_lidx_train(self.h, C.byref(p), C.byref(o)))
        d = {f: getattr(o, f) for f, _ in LidxInfo._fields_}
        d["root_loss"] = list(o.root_loss)
        d["leaf_loss"] = list(o.leaf_loss)
        return d

    def lidx_lookup(self, keys):
        keys = np.ascontiguousarray(keys, np.uint64)
        out = np.zeros(max(len(keys), 1), np.uint64)
        self._check(self.L.cf1g_lidx_lookup(self.h, _p(keys, U64P), len(keys), _p(out, U64P)))
        return out[:len(keys)]

    def lidx_export(self):
        sh = LidxShape()
        self._check(self.L.cf1g_lidx_export(self.h, C.byref(sh), None, None, None, None, None, None))
        keys = np.zeros(max(sh.n_keys, 1), np.uint64)
        warp = np.zeros((max(sh.warp_groups, 1), 256), np.uint32)
        root = np.zeros(sh.n_pieces + 1, ROOT_DTYPE)
        leaf = np.zeros(max(sh.n_leaves, 1), LEAF_DTYPE)
        err = np.zeros((max(sh.n_leaves, 1), 2), np.uint32)
        lines = np.zeros((max(sh.n_lines, 1), sh.line_bytes), np.uint8)
        self._check(self.L.cf1g_lidx_export(self.h, None, _p(keys, U64P), _p(warp, U32P), root.ctypes.data_as(C.c_void_p),
                                            leaf.ctypes.data_as(C.c_void_p), _p(err, U32P),
                                            lines.ctypes.data_as(C.c_void_p)))
        return {"keys": keys[:sh.n_keys], "warp": warp[:sh.warp_groups], "key_bits": sh.key_bits, "root": root, "leaf": leaf[:sh.n_leaves], "err": err[:sh.n_leaves],
                "lines": lines[:sh.n_lines], "shift": sh.root_shift, "line_bytes": sh.line_bytes,
                "line_keys": sh.line_keys}

    def lidx_gradients(self, level, w, c):
        w = np.ascontiguousarray(w, np.float32)
        c = np.ascontiguousarray(c, np.float32)
        n = len(w)
        gw, gc, sq = np.zeros(n), np.zeros(n), np.zeros(n)
        seg = np.zeros(n + 1, np.uint64)
        mu, isd = np.zeros(n, np.float32), np.zeros(n, np.float32)
        F32P = C.POINTER(C.c_float)
        F64P = C.POINTER(C.c_double)
        self._check(self.L.cf1g_lidx_gradients(self.h, level, _p(w, F32P), _p(c, F32P), n, _p(gw, F64P), _p(gc, F64P),
                                               _p(sq, F64P), _p(seg, U64P), _p(mu, F32P), _p(isd, F32P)))
        return {"gw": gw, "gc": gc, "sq": sq, "seg": seg, "mu": mu, "isd": isd}

    def classify_via(self, index):
        """P4 per occurrence (lookup + CAS) through an index; its CAS stats."""
        o = CasStats()
        self._check(self.L.cf1g_classify_via(self.h, index, C.byref(o)))
        return {f: getattr(o, f) for f, _ in CasStats._fields_}

    def select_index(self, index):
        self._check(self.L.cf1g_select_index(self.h, index))

    for f in (lidx_train, lidx_lookup, lidx_export, lidx_gradients, classify_via, select_index):
        setattr(Cf1g, f.__name__, f)


_lidx_methods()


def state_of(words, idx):
    """5-bit state idx from 12-per-word packed words (src/state_vector.rs:36-42)."""
    idx = np.asarray(idx, np.uint64)
    w = words[(idx // 12).astype(np.int64)]
    return ((w >> ((idx % 12) * 5)) & np.uint64(31)).astype(np.uint8)
