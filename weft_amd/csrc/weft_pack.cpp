// Host-side packing for the prepacked scoring drop-in (CPython extension _weft_pack).
//
// The reference launcher (src/weft/gpu_render.py:455-566) receives one record array and one
// count array per tile and packs eval descriptors with a Python loop per stack.  Here one
// pass over the Python lists (GIL held) collects pointers, lengths, tile offsets and sizes;
// the descriptors, the size-grouped launch order and the records themselves are then
// written by several threads with the GIL released, into one caller-provided (pinned)
// buffer laid out as
//
//     [E eval descriptors, 16 B][pad to 256][E int32 order][pad to 256][P records, 60 B]
//
// so that a single host-to-device copy moves all of it.  `split` returns the per-tile
// result views of one float32 array.  Every range a kernel would read is validated here,
// with the errors of the Python path (ValueError for bad ranges or sizes, KeyError for a
// missing tile offset); a layout this code does not handle (a record array of another
// dtype, a non-contiguous array, counts that are not an integer array) raises
// NotImplementedError so that the caller can take its NumPy path.
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int kSizes[4] = {8, 16, 32, 64};
constexpr int64_t kRec = 60;     // record bytes (WeftPrim)
constexpr int64_t kEval = 16;    // eval descriptor bytes (WeftEvalDesc)

inline int64_t align256(int64_t v) { return (v + 255) & ~int64_t(255); }

inline int size_class(long ts) {
    for (int k = 0; k < 4; ++k)
        if (ts == kSizes[k]) return k;
    return -1;
}

struct TileIn {
    const char* rec;     // record array data
    int64_t n_rec;       // records in the array
    const char* cnt;     // count array data
    int64_t n_cnt;       // stacks (evals) of this tile
    int cnt_type;        // numpy type number of the counts
    int64_t tile_off;    // float offset of the tile's pixels
    int ts;
    int cls;
    int64_t e_start;     // first eval of the tile
    int64_t r_start;     // first record of the tile in the packed array
    int64_t o_start;     // first position of the tile's evals in the order array
};

inline int64_t count_at(const TileIn& t, int64_t i) {
    switch (t.cnt_type) {
    case NPY_INT8: return reinterpret_cast<const int8_t*>(t.cnt)[i];
    case NPY_UINT8: return reinterpret_cast<const uint8_t*>(t.cnt)[i];
    case NPY_INT16: return reinterpret_cast<const int16_t*>(t.cnt)[i];
    case NPY_UINT16: return reinterpret_cast<const uint16_t*>(t.cnt)[i];
    case NPY_INT32: return reinterpret_cast<const int32_t*>(t.cnt)[i];
    case NPY_UINT32: return reinterpret_cast<const uint32_t*>(t.cnt)[i];
    case NPY_INT64: return reinterpret_cast<const int64_t*>(t.cnt)[i];
    default: return reinterpret_cast<const uint64_t*>(t.cnt)[i] > (uint64_t)INT64_MAX
                 ? -1 : (int64_t)reinterpret_cast<const uint64_t*>(t.cnt)[i];
    }
}

bool int_type_ok(int t) {
    return t == NPY_INT8 || t == NPY_UINT8 || t == NPY_INT16 || t == NPY_UINT16 || t == NPY_INT32 ||
           t == NPY_UINT32 || t == NPY_INT64 || t == NPY_UINT64 || t == NPY_LONGLONG || t == NPY_ULONGLONG ||
           t == NPY_LONG || t == NPY_ULONG;
}

int norm_int_type(PyArrayObject* a) {
    const int t = PyArray_TYPE(a);
    const int isz = (int)PyArray_ITEMSIZE(a);
    const bool sgn = PyTypeNum_ISSIGNED(t);
    switch (isz) {
    case 1: return sgn ? NPY_INT8 : NPY_UINT8;
    case 2: return sgn ? NPY_INT16 : NPY_UINT16;
    case 4: return sgn ? NPY_INT32 : NPY_UINT32;
    default: return sgn ? NPY_INT64 : NPY_UINT64;
    }
}

int n_workers(int requested) {
    int hw = (int)std::thread::hardware_concurrency();
    if (hw <= 0) hw = 1;
    int n = requested > 0 ? requested : hw;
    return std::max(1, std::min(n, 64));
}

// Run fn(worker, n_workers) on n threads (the calling thread is worker 0).
template <class F>
void run_threads(int n, F&& fn) {
    std::vector<std::thread> th;
    th.reserve(n > 0 ? n - 1 : 0);
    for (int w = 1; w < n; ++w) th.emplace_back([&fn, w, n] { fn(w, n); });
    fn(0, n);
    for (auto& t : th) t.join();
}

// Items [lo, hi) of worker w when `prefix` (size n+1, non-decreasing) weighs them.
void split_by_weight(const std::vector<int64_t>& prefix, int w, int nw, int64_t& lo, int64_t& hi) {
    const int64_t n = (int64_t)prefix.size() - 1;
    const int64_t total = prefix[n];
    auto first_at = [&](int64_t target) -> int64_t {
        return std::lower_bound(prefix.begin(), prefix.end(), target) - prefix.begin();
    };
    lo = w == 0 ? 0 : std::min<int64_t>(n, first_at((total * w + nw - 1) / nw));
    hi = w == nw - 1 ? n : std::min<int64_t>(n, first_at((total * (w + 1) + nw - 1) / nw));
    if (hi < lo) hi = lo;
}

PyArray_Descr* g_rec_descr = nullptr;  // the record dtype (set once from Python)

// pack(arrays, counts, tile_indices, tile_sizes, offsets, n_floats, out, threads)
//   -> ("need", nbytes) when `out` is too small, else
//      (n_evals, n_records, off_order, off_records, groups, e_start, n_per)
PyObject* py_pack(PyObject*, PyObject* args) {
    PyObject *arrays, *counts, *tidx, *tsizes, *offsets, *out;
    long long n_floats;
    int threads;
    if (!PyArg_ParseTuple(args, "OOOOOLOi", &arrays, &counts, &tidx, &tsizes, &offsets, &n_floats, &out, &threads))
        return nullptr;
    if (!g_rec_descr) {
        PyErr_SetString(PyExc_RuntimeError, "_weft_pack.set_record_dtype was not called");
        return nullptr;
    }
    PyObject* fa = PySequence_Fast(arrays, "eval_prim_arrays must be a sequence");
    if (!fa) return nullptr;
    PyObject* fc = PySequence_Fast(counts, "eval_prim_counts must be a sequence");
    PyObject* fi = fc ? PySequence_Fast(tidx, "tile_indices must be a sequence") : nullptr;
    PyObject* fs = fi ? PySequence_Fast(tsizes, "tile_sizes must be a sequence") : nullptr;
    std::vector<PyObject*> held;
    PyObject* result = nullptr;
    auto cleanup = [&] {
        for (PyObject* o : held) Py_DECREF(o);
        Py_XDECREF(fs);
        Py_XDECREF(fi);
        Py_XDECREF(fc);
        Py_DECREF(fa);
    };
    if (!fs) {
        cleanup();
        return nullptr;
    }
    const Py_ssize_t T = PySequence_Fast_GET_SIZE(fi);
    if (PySequence_Fast_GET_SIZE(fa) < T || PySequence_Fast_GET_SIZE(fc) < T || PySequence_Fast_GET_SIZE(fs) < T) {
        cleanup();
        PyErr_SetString(PyExc_IndexError, "per-tile lists are shorter than tile_indices");
        return nullptr;
    }
    std::vector<TileIn> tin((size_t)T);
    int64_t n_cls[4] = {0, 0, 0, 0};
    int64_t E = 0, P = 0;
    held.reserve(2 * (size_t)T);
    PyArray_Descr* last_ok = nullptr;
    for (Py_ssize_t j = 0; j < T; ++j) {
        PyObject* a = PySequence_Fast_GET_ITEM(fa, j);
        PyObject* c = PySequence_Fast_GET_ITEM(fc, j);
        if (!PyArray_Check(a) || !PyArray_Check(c)) {
            cleanup();
            PyErr_SetString(PyExc_NotImplementedError, "record or count list holds a non-array");
            return nullptr;
        }
        PyArrayObject* aa = reinterpret_cast<PyArrayObject*>(a);
        PyArrayObject* ca = reinterpret_cast<PyArrayObject*>(c);
        PyArray_Descr* d = PyArray_DESCR(aa);
        if (d != last_ok) {
            if (!PyArray_EquivTypes(d, g_rec_descr)) {
                cleanup();
                PyErr_SetString(PyExc_NotImplementedError, "record array of another dtype");
                return nullptr;
            }
            last_ok = d;
        }
        if (!PyArray_IS_C_CONTIGUOUS(aa) || !PyArray_IS_C_CONTIGUOUS(ca) || PyArray_NDIM(ca) > 1 ||
            !int_type_ok(PyArray_TYPE(ca)) || PyArray_ISBYTESWAPPED(ca)) {
            cleanup();
            PyErr_SetString(PyExc_NotImplementedError, "non-contiguous arrays or non-integer counts");
            return nullptr;
        }
        Py_INCREF(a);
        held.push_back(a);
        Py_INCREF(c);
        held.push_back(c);
        TileIn& t = tin[(size_t)j];
        t.rec = static_cast<const char*>(PyArray_DATA(aa));
        t.n_rec = (int64_t)PyArray_SIZE(aa);
        t.cnt = static_cast<const char*>(PyArray_DATA(ca));
        t.n_cnt = (int64_t)PyArray_SIZE(ca);
        t.cnt_type = norm_int_type(ca);
        const long ts = PyLong_AsLong(PySequence_Fast_GET_ITEM(fs, j));
        if (ts == -1 && PyErr_Occurred()) {
            cleanup();
            return nullptr;
        }
        t.ts = (int)ts;
        t.cls = size_class(ts);
        if (t.cls < 0) {
            cleanup();
            PyErr_SetString(PyExc_ValueError, "tile sizes must be in (8, 16, 32, 64)");
            return nullptr;
        }
        PyObject* off = PyObject_GetItem(offsets, PySequence_Fast_GET_ITEM(fi, j));
        if (!off) {
            cleanup();
            return nullptr;
        }
        const long long to = PyLong_AsLongLong(off);
        Py_DECREF(off);
        if (to == -1 && PyErr_Occurred()) {
            cleanup();
            return nullptr;
        }
        t.tile_off = (int64_t)to;
        if (t.n_cnt > 0 && (to < 0 || to + 3LL * ts * ts > n_floats)) {
            cleanup();
            PyErr_SetString(PyExc_ValueError, "eval tile ranges exceed the packed pixel buffer");
            return nullptr;
        }
        t.e_start = E;
        t.r_start = P;
        E += t.n_cnt;
        P += t.n_rec;
        n_cls[t.cls] += t.n_cnt;
    }
    if (P >= (int64_t)1 << 31 || E >= (int64_t)1 << 31 || n_floats >= (long long)1 << 31) {
        cleanup();
        PyErr_SetString(PyExc_ValueError, "batch too large for 32-bit offsets");
        return nullptr;
    }
    const int64_t off_ord = align256(E * kEval);
    const int64_t off_rec = align256(off_ord + 4 * E);
    const int64_t need = off_rec + std::max<int64_t>(P, 1) * kRec;
    Py_buffer ob;
    if (PyObject_GetBuffer(out, &ob, PyBUF_WRITABLE | PyBUF_C_CONTIGUOUS) < 0) {
        cleanup();
        return nullptr;
    }
    if (ob.len < need) {
        PyBuffer_Release(&ob);
        cleanup();
        return Py_BuildValue("(sL)", "need", (long long)need);
    }
    // launch order: evals grouped by tile size (8, 16, 32, 64), each group in eval order
    int64_t gbase[4], run[4];
    {
        int64_t s = 0;
        for (int k = 0; k < 4; ++k) {
            gbase[k] = s;
            run[k] = s;
            s += n_cls[k];
        }
        for (auto& t : tin) {
            t.o_start = run[t.cls];
            run[t.cls] += t.n_cnt;
        }
    }
    std::vector<int64_t> weight((size_t)T + 1, 0);
    for (Py_ssize_t j = 0; j < T; ++j)
        weight[(size_t)j + 1] = weight[(size_t)j] + tin[(size_t)j].n_rec * kRec + tin[(size_t)j].n_cnt * 20 + 64;
    char* base = static_cast<char*>(ob.buf);
    const int nw = (int)std::min<int64_t>(n_workers(threads), std::max<int64_t>(1, T / 64));
    std::vector<int> bad((size_t)nw, 0);
    Py_BEGIN_ALLOW_THREADS
    run_threads(nw, [&](int w, int n) {
        int64_t lo, hi;
        split_by_weight(weight, w, n, lo, hi);
        int32_t* ev = reinterpret_cast<int32_t*>(base);
        int32_t* ord = reinterpret_cast<int32_t*>(base + off_ord);
        for (int64_t j = lo; j < hi; ++j) {
            const TileIn& t = tin[(size_t)j];
            int64_t used = 0;
            for (int64_t i = 0; i < t.n_cnt; ++i) {
                const int64_t c = count_at(t, i);
                if (c < 0) {
                    bad[(size_t)w] = 1;
                    return;
                }
                int32_t* e = ev + 4 * (t.e_start + i);
                e[0] = (int32_t)(t.r_start + used);
                e[1] = (int32_t)c;
                e[2] = (int32_t)t.tile_off;
                e[3] = t.ts;
                ord[t.o_start + i] = (int32_t)(t.e_start + i);
                used += c;
                if (used > t.n_rec) {
                    bad[(size_t)w] = 1;
                    return;
                }
            }
            if (t.n_rec) std::memcpy(base + off_rec + t.r_start * kRec, t.rec, (size_t)(t.n_rec * kRec));
        }
    });
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&ob);
    for (int b : bad)
        if (b) {
            cleanup();
            PyErr_SetString(PyExc_ValueError, "eval primitive ranges exceed the packed primitive array");
            return nullptr;
        }
    npy_intp dims[1] = {(npy_intp)T};
    PyArrayObject* es = reinterpret_cast<PyArrayObject*>(PyArray_SimpleNew(1, dims, NPY_INT64));
    PyArrayObject* np_ = reinterpret_cast<PyArrayObject*>(PyArray_SimpleNew(1, dims, NPY_INT64));
    if (!es || !np_) {
        Py_XDECREF(es);
        Py_XDECREF(np_);
        cleanup();
        return nullptr;
    }
    int64_t* esd = static_cast<int64_t*>(PyArray_DATA(es));
    int64_t* npd = static_cast<int64_t*>(PyArray_DATA(np_));
    for (Py_ssize_t j = 0; j < T; ++j) {
        esd[j] = tin[(size_t)j].e_start;
        npd[j] = tin[(size_t)j].n_cnt;
    }
    PyObject* groups = PyList_New(0);
    for (int k = 0; k < 4 && groups; ++k) {
        if (!n_cls[k]) continue;
        PyObject* g = Py_BuildValue("(iL)", kSizes[k], (long long)n_cls[k]);
        if (!g || PyList_Append(groups, g) < 0) {
            Py_XDECREF(g);
            Py_CLEAR(groups);
            break;
        }
        Py_DECREF(g);
    }
    (void)gbase;
    if (groups)
        result = Py_BuildValue("(LLLLNNN)", (long long)E, (long long)P, (long long)off_ord, (long long)off_rec,
                               groups, reinterpret_cast<PyObject*>(es), reinterpret_cast<PyObject*>(np_));
    else {
        Py_DECREF(es);
        Py_DECREF(np_);
    }
    cleanup();
    return result;
}

// split(values, e_start, n_per) -> list of views values[e_start[j] : e_start[j] + n_per[j]]
PyObject* py_split(PyObject*, PyObject* args) {
    PyArrayObject *vals, *es, *np_;
    if (!PyArg_ParseTuple(args, "O!O!O!", &PyArray_Type, &vals, &PyArray_Type, &es, &PyArray_Type, &np_))
        return nullptr;
    if (PyArray_NDIM(vals) != 1 || !PyArray_IS_C_CONTIGUOUS(vals) || PyArray_TYPE(es) != NPY_INT64 ||
        PyArray_TYPE(np_) != NPY_INT64 || PyArray_SIZE(es) != PyArray_SIZE(np_) || !PyArray_IS_C_CONTIGUOUS(es) ||
        !PyArray_IS_C_CONTIGUOUS(np_)) {
        PyErr_SetString(PyExc_ValueError, "split: bad arguments");
        return nullptr;
    }
    const npy_intp T = PyArray_SIZE(es);
    const npy_intp n = PyArray_SIZE(vals);
    const int64_t* e = static_cast<const int64_t*>(PyArray_DATA(es));
    const int64_t* c = static_cast<const int64_t*>(PyArray_DATA(np_));
    for (npy_intp j = 0; j < T; ++j)
        if (e[j] < 0 || c[j] < 0 || e[j] + c[j] > n) {
            PyErr_SetString(PyExc_ValueError, "split: range outside the values");
            return nullptr;
        }
    PyObject* list = PyList_New(T);
    if (!list) return nullptr;
    PyArray_Descr* d = PyArray_DESCR(vals);
    const npy_intp isz = PyArray_ITEMSIZE(vals);
    char* data = static_cast<char*>(PyArray_DATA(vals));
    const int flags = PyArray_FLAGS(vals) & (NPY_ARRAY_WRITEABLE | NPY_ARRAY_ALIGNED);
    for (npy_intp j = 0; j < T; ++j) {
        npy_intp dims[1] = {(npy_intp)c[j]};
        Py_INCREF(d);
        PyObject* v = PyArray_NewFromDescr(&PyArray_Type, d, 1, dims, nullptr, data + e[j] * isz,
                                           flags | NPY_ARRAY_C_CONTIGUOUS, nullptr);
        if (!v) {
            Py_DECREF(list);
            return nullptr;
        }
        Py_INCREF(vals);
        if (PyArray_SetBaseObject(reinterpret_cast<PyArrayObject*>(v), reinterpret_cast<PyObject*>(vals)) < 0) {
            Py_DECREF(v);
            Py_DECREF(list);
            return nullptr;
        }
        PyList_SET_ITEM(list, j, v);
    }
    return list;
}

// equal(a, b, threads) -> bool: byte equality of two buffers of the same length, compared
// by several threads.
PyObject* py_equal(PyObject*, PyObject* args) {
    PyObject *a, *b;
    int threads;
    if (!PyArg_ParseTuple(args, "OOi", &a, &b, &threads)) return nullptr;
    Py_buffer va, vb;
    if (PyObject_GetBuffer(a, &va, PyBUF_C_CONTIGUOUS) < 0) return nullptr;
    if (PyObject_GetBuffer(b, &vb, PyBUF_C_CONTIGUOUS) < 0) {
        PyBuffer_Release(&va);
        return nullptr;
    }
    bool same = va.len == vb.len;
    if (same && va.len > 0) {
        const char* pa = static_cast<const char*>(va.buf);
        const char* pb = static_cast<const char*>(vb.buf);
        const int64_t len = va.len;
        const int nw = (int)std::min<int64_t>(n_workers(threads), std::max<int64_t>(1, len >> 20));
        std::vector<int> diff((size_t)nw, 0);
        Py_BEGIN_ALLOW_THREADS
        run_threads(nw, [&](int w, int n) {
            const int64_t lo = len * w / n, hi = len * (w + 1) / n;
            diff[(size_t)w] = std::memcmp(pa + lo, pb + lo, (size_t)(hi - lo)) != 0;
        });
        Py_END_ALLOW_THREADS
        for (int d : diff) same = same && !d;
    }
    PyBuffer_Release(&vb);
    PyBuffer_Release(&va);
    return PyBool_FromLong(same);
}

// copy(dst, src, threads): dst[:] = src (same byte length), by several threads.
PyObject* py_copy(PyObject*, PyObject* args) {
    PyObject *a, *b;
    int threads;
    if (!PyArg_ParseTuple(args, "OOi", &a, &b, &threads)) return nullptr;
    Py_buffer vd, vs;
    if (PyObject_GetBuffer(a, &vd, PyBUF_WRITABLE | PyBUF_C_CONTIGUOUS) < 0) return nullptr;
    if (PyObject_GetBuffer(b, &vs, PyBUF_C_CONTIGUOUS) < 0) {
        PyBuffer_Release(&vd);
        return nullptr;
    }
    if (vd.len != vs.len) {
        PyBuffer_Release(&vs);
        PyBuffer_Release(&vd);
        PyErr_SetString(PyExc_ValueError, "copy: lengths differ");
        return nullptr;
    }
    char* pd = static_cast<char*>(vd.buf);
    const char* ps = static_cast<const char*>(vs.buf);
    const int64_t len = vd.len;
    const int nw = (int)std::min<int64_t>(n_workers(threads), std::max<int64_t>(1, len >> 20));
    Py_BEGIN_ALLOW_THREADS
    run_threads(nw, [&](int w, int n) {
        const int64_t lo = len * w / n, hi = len * (w + 1) / n;
        std::memcpy(pd + lo, ps + lo, (size_t)(hi - lo));
    });
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&vs);
    PyBuffer_Release(&vd);
    Py_RETURN_NONE;
}

PyObject* py_set_record_dtype(PyObject*, PyObject* args) {
    PyObject* d;
    if (!PyArg_ParseTuple(args, "O", &d)) return nullptr;
    PyArray_Descr* descr = nullptr;
    if (!PyArray_DescrConverter(d, &descr)) return nullptr;
    if (PyDataType_ELSIZE(descr) != kRec) {
        Py_DECREF(descr);
        PyErr_SetString(PyExc_ValueError, "record dtype must be 60 bytes");
        return nullptr;
    }
    Py_XDECREF(g_rec_descr);
    g_rec_descr = descr;
    Py_RETURN_NONE;
}

PyMethodDef kMethods[] = {
    {"pack", py_pack, METH_VARARGS, "Pack eval descriptors, launch order and records into one buffer."},
    {"split", py_split, METH_VARARGS, "Per-tile views of one result array."},
    {"equal", py_equal, METH_VARARGS, "Byte equality of two buffers (threaded)."},
    {"copy", py_copy, METH_VARARGS, "Threaded byte copy between two buffers of one length."},
    {"set_record_dtype", py_set_record_dtype, METH_VARARGS, "The 60-byte record dtype."},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_weft_pack", "Host-side packing for the prepacked scorer.", -1,
                       kMethods};

}  // namespace

PyMODINIT_FUNC PyInit__weft_pack(void) {
    import_array();
    return PyModule_Create(&kModule);
}
