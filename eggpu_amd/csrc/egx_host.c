/* egx_host.c -- host side of the dict-graph route: the adjacency of an
 * EasyGraph-protocol graph (a dict of node -> dict of neighbour -> edge data)
 * read straight into int32 / float64 arrays, at C speed, for the device
 * build.  CPython C API only (buffers through the buffer protocol).
 *
 *   arcs(adj, index, n, drop_self, V, E, weight, W) -> m
 *     Row i is the i-th key of adj (insertion order, graph_convert.cpp:46-58);
 *     its arcs are the keys of adj[u] in their order, self-loops skipped
 *     when drop_self (undirected graphs, graph_convert.cpp:72-108: every
 *     pair then appears once in each endpoint's row, which is the
 *     both-directions out-CSR), each mapped to its row through `index`
 *     (a dict node -> position) or, when index is None, taken as the int it
 *     is (nodes 0..n-1 in order).  V (int32[n+1]) receives the row offsets,
 *     E (int32[capacity]) the columns; when W (float64[capacity]) is not
 *     None it receives float(data.get(weight, 1.0)) per arc.  Returns the
 *     arc count; raises if a neighbour is unknown or the capacity is short.
 *   sort_rows(V, E, W) -> rows_sorted
 *     Sorts every row's columns ascending (W along), stably, in place, as
 *     Graph::gen_CSR does (graph.cpp:855-878); rows already in order are
 *     only checked.
 *   identity(nodes) -> bool
 *     True when the list's i-th item is the int i for every i (exact int
 *     type), so that a node is its own row index.
 *   stamp(adj) -> (fingerprint, arcs)
 *     A 64-bit fingerprint of the node keys' hashes, their order, their
 *     degrees and their first neighbours' hashes (splitmix64 finaliser over
 *     each position, summed and xored), and the number of adjacency entries.
 *   int_arcs(adj, n, drop_self, threads) -> (V, E, m, fingerprint, arcs) | None
 *     arcs() and stamp() in one pass for the common case: the adjacency's
 *     keys are the ints 0..n-1 in order, every neighbour key is an int and
 *     there are no weights.  V and E are bytearrays of int32 (E holds at
 *     least m columns); the fingerprint and arc count are stamp()'s.  The
 *     rows are split over up to `threads` POSIX threads by arc count.  The
 *     calling thread keeps the GIL throughout, so no Python code runs, and
 *     no dict changes, while the workers read; they only read dict entries
 *     and int values (no references taken, no allocation, no Python
 *     errors).  None when a key falls outside that case: the caller then
 *     uses arcs() and stamp(), which raise the proper errors.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int get_buf(PyObject* obj, Py_buffer* b, Py_ssize_t itemsize, const char* what) {
    if (PyObject_GetBuffer(obj, b, PyBUF_WRITABLE | PyBUF_C_CONTIGUOUS) != 0) return -1;
    if (b->itemsize != itemsize) {
        PyErr_Format(PyExc_TypeError, "%s: expected %zd-byte items", what, itemsize);
        PyBuffer_Release(b);
        return -1;
    }
    return 0;
}

static PyObject* arcs(PyObject* self, PyObject* args) {
    (void)self;
    PyObject *adj, *index, *Vo, *Eo, *weight, *Wo;
    Py_ssize_t n;
    int drop_self;
    if (!PyArg_ParseTuple(args, "OOnpOOOO", &adj, &index, &n, &drop_self, &Vo, &Eo, &weight, &Wo)) return NULL;
    if (!PyDict_Check(adj) || (index != Py_None && !PyDict_Check(index))) {
        PyErr_SetString(PyExc_TypeError, "arcs: adj and index must be dicts");
        return NULL;
    }
    Py_buffer vb, eb, wb;
    if (get_buf(Vo, &vb, 4, "V") != 0) return NULL;
    if (get_buf(Eo, &eb, 4, "E") != 0) {
        PyBuffer_Release(&vb);
        return NULL;
    }
    const int with_w = Wo != Py_None;
    if (with_w && get_buf(Wo, &wb, 8, "W") != 0) {
        PyBuffer_Release(&vb);
        PyBuffer_Release(&eb);
        return NULL;
    }
    int32_t* V = (int32_t*)vb.buf;
    int32_t* E = (int32_t*)eb.buf;
    double* W = with_w ? (double*)wb.buf : NULL;
    const Py_ssize_t cap = eb.len / 4;
    PyObject* err = NULL;
    Py_ssize_t row = 0, m = 0, pos = 0;
    PyObject *u, *nbrs;
    if (vb.len / 4 < n + 1 || (with_w && wb.len / 8 < cap)) {
        err = PyExc_ValueError;
        PyErr_SetString(err, "arcs: output arrays too small");
        goto done;
    }
    V[0] = 0;
    while (PyDict_Next(adj, &pos, &u, &nbrs)) {
        if (row >= n) {
            err = PyExc_ValueError;
            PyErr_SetString(err, "arcs: adjacency has more rows than nodes");
            goto done;
        }
        if (!PyDict_Check(nbrs)) {
            err = PyExc_TypeError;
            PyErr_SetString(err, "arcs: adjacency rows must be dicts");
            goto done;
        }
        Py_ssize_t q = 0;
        PyObject *v, *data;
        while (PyDict_Next(nbrs, &q, &v, &data)) {
            long col;
            if (index == Py_None) {
                col = PyLong_AsLong(v);
                if (col == -1 && PyErr_Occurred()) {
                    err = PyExc_TypeError;
                    goto done;
                }
            } else {
                PyObject* ix = PyDict_GetItemWithError(index, v);
                if (!ix) {
                    err = PyExc_KeyError;
                    if (!PyErr_Occurred()) PyErr_SetObject(PyExc_KeyError, v);
                    goto done;
                }
                col = PyLong_AsLong(ix);
            }
            if (col < 0 || col >= n) {
                err = PyExc_ValueError;
                PyErr_SetString(err, "arcs: neighbour outside the node range");
                goto done;
            }
            if (drop_self && col == row) continue;
            if (m >= cap) {
                err = PyExc_ValueError;
                PyErr_SetString(err, "arcs: more arcs than the capacity");
                goto done;
            }
            E[m] = (int32_t)col;
            if (with_w) {
                double w = 1.0;
                if (PyDict_Check(data)) {
                    PyObject* x = PyDict_GetItemWithError(data, weight);
                    if (x) {
                        w = PyFloat_AsDouble(x);
                        if (w == -1.0 && PyErr_Occurred()) {
                            err = PyExc_TypeError;
                            goto done;
                        }
                    } else if (PyErr_Occurred()) {
                        err = PyExc_TypeError;
                        goto done;
                    }
                }
                W[m] = w;
            }
            ++m;
        }
        if (m > INT32_MAX - 1) {
            err = PyExc_OverflowError;
            PyErr_SetString(err, "arcs: arc count exceeds the signed 32-bit CSR limit");
            goto done;
        }
        V[++row] = (int32_t)m;
    }
    if (row != n) {
        err = PyExc_ValueError;
        PyErr_SetString(err, "arcs: adjacency has fewer rows than nodes");
    }
done:
    PyBuffer_Release(&vb);
    PyBuffer_Release(&eb);
    if (with_w) PyBuffer_Release(&wb);
    if (err) return NULL;
    return PyLong_FromSsize_t(m);
}

/* stable insertion sort of one short row; a merge sort for long ones */
static void sort_row(int32_t* e, double* w, Py_ssize_t len, int32_t* te, double* tw) {
    if (len <= 32) {
        for (Py_ssize_t i = 1; i < len; ++i) {
            const int32_t k = e[i];
            const double kw = w ? w[i] : 0.0;
            Py_ssize_t j = i;
            while (j > 0 && e[j - 1] > k) {
                e[j] = e[j - 1];
                if (w) w[j] = w[j - 1];
                --j;
            }
            e[j] = k;
            if (w) w[j] = kw;
        }
        return;
    }
    const Py_ssize_t h = len / 2;
    sort_row(e, w, h, te, tw);
    sort_row(e + h, w ? w + h : NULL, len - h, te, tw);
    Py_ssize_t a = 0, b = h, o = 0;
    while (a < h && b < len) {
        if (e[b] < e[a]) {
            te[o] = e[b];
            if (w) tw[o] = w[b];
            ++b;
        } else {
            te[o] = e[a];
            if (w) tw[o] = w[a];
            ++a;
        }
        ++o;
    }
    while (a < h) {
        te[o] = e[a];
        if (w) tw[o] = w[a];
        ++a, ++o;
    }
    while (b < len) {
        te[o] = e[b];
        if (w) tw[o] = w[b];
        ++b, ++o;
    }
    memcpy(e, te, (size_t)len * 4);
    if (w) memcpy(w, tw, (size_t)len * 8);
}

static PyObject* sort_rows(PyObject* self, PyObject* args) {
    (void)self;
    PyObject *Vo, *Eo, *Wo;
    if (!PyArg_ParseTuple(args, "OOO", &Vo, &Eo, &Wo)) return NULL;
    Py_buffer vb, eb, wb;
    if (get_buf(Vo, &vb, 4, "V") != 0) return NULL;
    if (get_buf(Eo, &eb, 4, "E") != 0) {
        PyBuffer_Release(&vb);
        return NULL;
    }
    const int with_w = Wo != Py_None;
    if (with_w && get_buf(Wo, &wb, 8, "W") != 0) {
        PyBuffer_Release(&vb);
        PyBuffer_Release(&eb);
        return NULL;
    }
    const int32_t* V = (const int32_t*)vb.buf;
    int32_t* E = (int32_t*)eb.buf;
    double* W = with_w ? (double*)wb.buf : NULL;
    const Py_ssize_t n = vb.len / 4 - 1;
    Py_ssize_t sorted_rows = 0, longest = 0;
    int ok = 1;
    int32_t* te = NULL;
    double* tw = NULL;
    for (Py_ssize_t r = 0; r < n && ok; ++r) {
        if (V[r] < 0 || V[r] > V[r + 1] || (Py_ssize_t)V[r + 1] > eb.len / 4) ok = 0;
        else if (V[r + 1] - V[r] > longest) longest = V[r + 1] - V[r];
    }
    if (!ok) {
        PyErr_SetString(PyExc_ValueError, "sort_rows: malformed offsets");
        goto out;
    }
    te = (int32_t*)malloc((size_t)(longest + 1) * 4);
    tw = with_w ? (double*)malloc((size_t)(longest + 1) * 8) : NULL;
    if (!te || (with_w && !tw)) {
        free(te);
        free(tw);
        PyErr_NoMemory();
        ok = 0;
        goto out;
    }
    for (Py_ssize_t r = 0; r < n; ++r) {
        const Py_ssize_t b = V[r], e = V[r + 1];
        Py_ssize_t i = b + 1;
        while (i < e && E[i - 1] <= E[i]) ++i;
        if (i >= e) continue;
        sort_row(E + b, W ? W + b : NULL, e - b, te, tw);
        ++sorted_rows;
    }
    free(te);
    free(tw);
out:
    PyBuffer_Release(&vb);
    PyBuffer_Release(&eb);
    if (with_w) PyBuffer_Release(&wb);
    if (!ok) return NULL;
    return PyLong_FromSsize_t(sorted_rows);
}

static uint64_t mix(uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

/* the per-position term of stamp()'s fingerprint */
static uint64_t stamp_term(uint64_t h, uint64_t d, uint64_t i, uint64_t h1) {
    return mix(h * 0x9E3779B97F4A7C15ull + d * 0xD6E8FEB86659FD93ull + i * 0xA0761D6478BD642Full +
               h1 * 0xE7037ED1A0B428DBull);
}

static PyObject* stamp(PyObject* self, PyObject* args) {
    (void)self;
    PyObject* adj;
    if (!PyArg_ParseTuple(args, "O", &adj)) return NULL;
    if (!PyDict_Check(adj)) {
        PyErr_SetString(PyExc_TypeError, "stamp: adj must be a dict");
        return NULL;
    }
    uint64_t sum = 0, xr = 0x243F6A8885A308D3ull;
    Py_ssize_t pos = 0, i = 0, arcs_total = 0;
    PyObject *u, *nbrs;
    while (PyDict_Next(adj, &pos, &u, &nbrs)) {
        const Py_hash_t h = PyObject_Hash(u);
        if (h == -1 && PyErr_Occurred()) return NULL;
        if (!PyDict_Check(nbrs)) {
            PyErr_SetString(PyExc_TypeError, "stamp: adjacency rows must be dicts");
            return NULL;
        }
        const Py_ssize_t d = PyDict_GET_SIZE(nbrs);
        Py_hash_t h1 = 0;
        Py_ssize_t q = 0;
        PyObject *v, *data;
        if (PyDict_Next(nbrs, &q, &v, &data)) {
            h1 = PyObject_Hash(v);
            if (h1 == -1 && PyErr_Occurred()) return NULL;
        }
        const uint64_t x = stamp_term((uint64_t)h, (uint64_t)d, (uint64_t)i, (uint64_t)h1);
        sum += x;
        xr ^= mix(x + (uint64_t)i);
        arcs_total += d;
        ++i;
    }
    const uint64_t fp = mix(sum ^ (xr * 0x9E3779B97F4A7C15ull)) ^ (uint64_t)i;
    return Py_BuildValue("(Kn)", (unsigned long long)fp, arcs_total);
}

typedef struct {
    PyObject* const* rows; /* the neighbour dict of each row (borrowed) */
    const int64_t* off;    /* row i's columns go to E[off[i]...] */
    int32_t* E;
    int32_t* kept;         /* columns kept per row */
    Py_ssize_t r0, r1, n;
    int drop_self;
    uint64_t sum, xr;
    int bad;
} IntArcsPart;

static void* int_arcs_part(void* arg) {
    IntArcsPart* p = (IntArcsPart*)arg;
    for (Py_ssize_t i = p->r0; i < p->r1; ++i) {
        int32_t* e = p->E + p->off[i];
        Py_ssize_t q = 0, k = 0;
        PyObject *v, *data;
        uint64_t h1 = 0;
        int first = 1;
        while (PyDict_Next(p->rows[i], &q, &v, &data)) {
            if (!PyLong_CheckExact(v)) {
                p->bad = 1;
                return NULL;
            }
            int of = 0;
            const long long c = PyLong_AsLongLongAndOverflow(v, &of);
            if (of || c < 0 || c >= p->n) {
                p->bad = 1;
                return NULL;
            }
            if (first) h1 = (uint64_t)c; /* hash(c) == c for 0 <= c < 2^61 - 1 */
            first = 0;
            if (p->drop_self && c == i) continue;
            e[k++] = (int32_t)c;
        }
        p->kept[i] = (int32_t)k;
        const uint64_t x = stamp_term((uint64_t)i, (uint64_t)(p->off[i + 1] - p->off[i]), (uint64_t)i, h1);
        p->sum += x;
        p->xr ^= mix(x + (uint64_t)i);
    }
    return NULL;
}

#define INT_ARCS_MAX_THREADS 32

static PyObject* int_arcs(PyObject* self, PyObject* args) {
    (void)self;
    PyObject* adj;
    Py_ssize_t n;
    int drop_self, threads;
    if (!PyArg_ParseTuple(args, "Onpi", &adj, &n, &drop_self, &threads)) return NULL;
    if (!PyDict_Check(adj) || n < 0 || PyDict_GET_SIZE(adj) != n) Py_RETURN_NONE;
    PyObject** rows = (PyObject**)malloc(sizeof(PyObject*) * (size_t)(n > 0 ? n : 1));
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t* kept = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    PyObject *Vb = NULL, *Eb = NULL, *res = NULL;
    if (!rows || !off || !kept) {
        PyErr_NoMemory();
        goto out;
    }
    /* rows in adjacency order, each keyed by its own index */
    Py_ssize_t pos = 0, i = 0;
    PyObject *u, *nbrs;
    off[0] = 0;
    while (PyDict_Next(adj, &pos, &u, &nbrs)) {
        int of = 0;
        if (i >= n || !PyLong_CheckExact(u) || PyLong_AsLongLongAndOverflow(u, &of) != (long long)i || of ||
            !PyDict_Check(nbrs)) {
            res = Py_None;
            goto out;
        }
        rows[i] = nbrs;
        off[i + 1] = off[i] + PyDict_GET_SIZE(nbrs);
        ++i;
    }
    const int64_t total = off[n];
    if (i != n || total > INT32_MAX - 1) {
        res = Py_None;
        goto out;
    }
    Vb = PyByteArray_FromStringAndSize(NULL, (Py_ssize_t)(n + 1) * 4);
    Eb = PyByteArray_FromStringAndSize(NULL, (Py_ssize_t)(total > 0 ? total : 1) * 4);
    if (!Vb || !Eb) goto out;
    int32_t* V = (int32_t*)PyByteArray_AS_STRING(Vb);
    int32_t* E = (int32_t*)PyByteArray_AS_STRING(Eb);
    int T = threads < 1 ? 1 : (threads > INT_ARCS_MAX_THREADS ? INT_ARCS_MAX_THREADS : threads);
    if (total < 65536 || n < 2 * T) T = 1;
    IntArcsPart parts[INT_ARCS_MAX_THREADS];
    pthread_t tid[INT_ARCS_MAX_THREADS];
    int started[INT_ARCS_MAX_THREADS];
    Py_ssize_t r = 0;
    for (int t = 0; t < T; ++t) {
        /* rows [r, r1): the first row whose arcs start at or past (t + 1) / T of them */
        const int64_t goal = total * (t + 1) / T;
        Py_ssize_t r1 = r;
        if (t == T - 1) r1 = n;
        else
            while (r1 < n && off[r1] < goal) ++r1;
        parts[t] = (IntArcsPart){rows, off, E, kept, r, r1, n, drop_self, 0, 0, 0};
        r = r1;
    }
    for (int t = 1; t < T; ++t) started[t] = pthread_create(&tid[t], NULL, int_arcs_part, &parts[t]) == 0;
    int_arcs_part(&parts[0]);
    for (int t = 1; t < T; ++t) {
        if (started[t]) pthread_join(tid[t], NULL);
        else int_arcs_part(&parts[t]); /* no thread: this one runs it */
    }
    uint64_t sum = 0, xr = 0x243F6A8885A308D3ull;
    for (int t = 0; t < T; ++t) {
        if (parts[t].bad) {
            res = Py_None;
            goto out;
        }
        sum += parts[t].sum;
        xr ^= parts[t].xr;
    }
    /* rows that dropped a self-loop leave gaps: close them, in order */
    int64_t m = 0;
    V[0] = 0;
    for (Py_ssize_t k = 0; k < n; ++k) {
        if (m != off[k] && kept[k] > 0) memmove(E + m, E + off[k], (size_t)kept[k] * 4);
        m += kept[k];
        V[k + 1] = (int32_t)m;
    }
    const uint64_t fp = mix(sum ^ (xr * 0x9E3779B97F4A7C15ull)) ^ (uint64_t)n;
    res = Py_BuildValue("(OOnKn)", Vb, Eb, (Py_ssize_t)m, (unsigned long long)fp, (Py_ssize_t)total);
out:
    free(rows);
    free(off);
    free(kept);
    Py_XDECREF(Vb);
    Py_XDECREF(Eb);
    if (res == Py_None) Py_RETURN_NONE;
    return res;
}

/* identity(nodes) -> bool: nodes is a list whose i-th item is the int i
 * (exact int type, bool excluded), so that every node is its own index. */
static PyObject* identity(PyObject* self, PyObject* args) {
    (void)self;
    PyObject* nodes;
    if (!PyArg_ParseTuple(args, "O", &nodes)) return NULL;
    if (!PyList_Check(nodes)) {
        PyErr_SetString(PyExc_TypeError, "identity: nodes must be a list");
        return NULL;
    }
    const Py_ssize_t n = PyList_GET_SIZE(nodes);
    for (Py_ssize_t i = 0; i < n; ++i) {
        PyObject* v = PyList_GET_ITEM(nodes, i);
        if (!PyLong_CheckExact(v)) Py_RETURN_FALSE;
        int overflow = 0;
        const long long x = PyLong_AsLongLongAndOverflow(v, &overflow);
        if (overflow || x != (long long)i) Py_RETURN_FALSE;
    }
    Py_RETURN_TRUE;
}

static PyMethodDef methods[] = {
    {"identity", identity, METH_VARARGS, "nodes are exactly the ints 0..n-1 in order"},
    {"arcs", arcs, METH_VARARGS, "adjacency dicts -> out-CSR arrays"},
    {"sort_rows", sort_rows, METH_VARARGS, "sort each CSR row's columns ascending, in place"},
    {"stamp", stamp, METH_VARARGS, "fingerprint of node keys, order, degrees, first neighbours; arc count"},
    {"int_arcs", int_arcs, METH_VARARGS, "arcs() and stamp() in one threaded pass for int-keyed, unweighted adjacencies"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_egx_host", NULL, -1, methods, NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__egx_host(void) { return PyModule_Create(&module); }
