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
 *     rows are split over up to `threads` threads (a pool kept for the
 *     process) by arc count.  The
 *     calling thread keeps the GIL throughout, so no Python code runs, and
 *     no dict changes, while the workers read; they only read dict entries
 *     and int values (no references taken, no allocation, no Python
 *     errors).  None when a key falls outside that case: the caller then
 *     uses arcs() and stamp(), which raise the proper errors.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <pthread.h>
#include <unistd.h>
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

#define INT_ARCS_MAX_THREADS 32

/* A pool of worker threads for int_arcs, made on first use and kept for the
 * process (creating a thread costs tens of microseconds, a cold R-MAT 14
 * walk a few hundred).  run_parallel(T, fn, args, size) runs fn on args[0]
 * here and on args[1..T-1] in pool threads, and returns once all are done.
 * A process forked from this one starts a pool of its own. */
typedef void (*PoolFn)(void*);
typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t go, done;
    pid_t pid;
    int workers;      /* threads started */
    unsigned long gen; /* job generation */
    PoolFn fn;
    char* args;
    size_t size;
    int jobs;         /* workers taking part in the current job */
    int left;         /* of those, still running */
} WalkPool;

static WalkPool walk_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, 0,
                             NULL, NULL, 0, 0, 0};

typedef struct {
    int id;              /* 1.. */
    unsigned long gen0;  /* the job generation when the thread was made */
} PoolSeat;

static PoolSeat pool_seats[INT_ARCS_MAX_THREADS];

static void* pool_worker(void* arg) {
    const int id = ((PoolSeat*)arg)->id;
    WalkPool* P = &walk_pool;
    pthread_mutex_lock(&P->mu);
    /* a job posted after this thread was made is one it takes part in */
    unsigned long seen = ((PoolSeat*)arg)->gen0;
    for (;;) {
        while (P->gen == seen) pthread_cond_wait(&P->go, &P->mu);
        seen = P->gen;
        if (id >= P->jobs) continue;
        PoolFn fn = P->fn;
        void* a = P->args + (size_t)id * P->size;
        pthread_mutex_unlock(&P->mu);
        fn(a);
        pthread_mutex_lock(&P->mu);
        if (--P->left == 0) pthread_cond_signal(&P->done);
    }
    return NULL;
}

/* the number of pool threads that can take part (at most want - 1) */
static int pool_ready(int want) {
    WalkPool* P = &walk_pool;
    if (P->pid != getpid()) { /* first use, or a forked child: no threads here */
        pthread_mutex_init(&P->mu, NULL);
        pthread_cond_init(&P->go, NULL);
        pthread_cond_init(&P->done, NULL);
        P->pid = getpid();
        P->workers = 0;
        P->gen = 0;
    }
    while (P->workers < want - 1 && P->workers < INT_ARCS_MAX_THREADS - 1) {
        pthread_t t;
        pthread_attr_t at;
        pool_seats[P->workers].id = P->workers + 1;
        pool_seats[P->workers].gen0 = P->gen; /* gen changes only in run_parallel, on this thread */
        pthread_attr_init(&at);
        pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
        const int rc = pthread_create(&t, &at, pool_worker, &pool_seats[P->workers]);
        pthread_attr_destroy(&at);
        if (rc != 0) break;
        ++P->workers;
    }
    return P->workers;
}

static void run_parallel(int T, PoolFn fn, void* args, size_t size) {
    WalkPool* P = &walk_pool;
    if (T > 1) {
        pthread_mutex_lock(&P->mu);
        P->fn = fn;
        P->args = (char*)args;
        P->size = size;
        P->jobs = T;
        P->left = T - 1;
        ++P->gen;
        pthread_cond_broadcast(&P->go);
        pthread_mutex_unlock(&P->mu);
    }
    fn(args);
    if (T > 1) {
        pthread_mutex_lock(&P->mu);
        while (P->left > 0) pthread_cond_wait(&P->done, &P->mu);
        pthread_mutex_unlock(&P->mu);
    }
}

/* What the parts of one int_arcs call share.  Phase 1 (rows split evenly):
 * check each row's key and neighbour dict and take its size.  The calling
 * thread then turns the sizes into offsets, allocates E and splits the rows
 * by arc count.  Phase 2: read each row's neighbours into E. */
typedef struct {
    PyObject* const* keys; /* the adjacency's keys and neighbour dicts, borrowed */
    PyObject* const* rows;
    int64_t* off;          /* row i's columns go to E[off[i]...] */
    int32_t* E;
    int32_t* kept;         /* columns kept per row */
    Py_ssize_t n;
    int drop_self, T;
    Py_ssize_t cut[INT_ARCS_MAX_THREADS + 1]; /* phase 2 row ranges */
} IntArcsJob;

typedef struct {
    IntArcsJob* job;
    int t;
    uint64_t sum, xr;
    int bad;
} IntArcsPart;

static void int_arcs_sizes(void* arg) {
    IntArcsPart* p = (IntArcsPart*)arg;
    IntArcsJob* j = p->job;
    const Py_ssize_t r0 = j->n * p->t / j->T, r1 = j->n * (p->t + 1) / j->T;
    for (Py_ssize_t i = r0; i < r1; ++i) {
        PyObject* u = j->keys[i];
        int of = 0;
        if (!PyLong_CheckExact(u) || PyLong_AsLongLongAndOverflow(u, &of) != (long long)i || of ||
            !PyDict_Check(j->rows[i])) {
            p->bad = 1;
            return;
        }
        j->off[i + 1] = PyDict_GET_SIZE(j->rows[i]);
    }
}

static void int_arcs_fill(void* arg) {
    IntArcsPart* p = (IntArcsPart*)arg;
    IntArcsJob* j = p->job;
    for (Py_ssize_t i = j->cut[p->t]; i < j->cut[p->t + 1]; ++i) {
        int32_t* e = j->E + j->off[i];
        Py_ssize_t q = 0, k = 0;
        PyObject *v, *data;
        uint64_t h1 = 0;
        int first = 1;
        while (PyDict_Next(j->rows[i], &q, &v, &data)) {
            if (!PyLong_CheckExact(v)) {
                p->bad = 1;
                return;
            }
            int of = 0;
            const long long c = PyLong_AsLongLongAndOverflow(v, &of);
            if (of || c < 0 || c >= j->n) {
                p->bad = 1;
                return;
            }
            if (first) h1 = (uint64_t)c; /* hash(c) == c for 0 <= c < 2^61 - 1 */
            first = 0;
            if (j->drop_self && c == i) continue;
            e[k++] = (int32_t)c;
        }
        j->kept[i] = (int32_t)k;
        const uint64_t x = stamp_term((uint64_t)i, (uint64_t)(j->off[i + 1] - j->off[i]), (uint64_t)i, h1);
        p->sum += x;
        p->xr ^= mix(x + (uint64_t)i);
    }
}

/* between the phases: offsets, E, the split by arcs; returns 0, 1 (a key
 * outside the case) or 2 (allocation failed, exception set) */
static int int_arcs_setup(IntArcsJob* j, IntArcsPart* parts, PyObject** Vb, PyObject** Eb, int64_t* total) {
    for (int t = 0; t < j->T; ++t)
        if (parts[t].bad) return 1;
    const Py_ssize_t n = j->n;
    j->off[0] = 0;
    for (Py_ssize_t k = 0; k < n; ++k) j->off[k + 1] += j->off[k];
    *total = j->off[n];
    if (*total > INT32_MAX - 1) return 1;
    *Vb = PyByteArray_FromStringAndSize(NULL, (Py_ssize_t)(n + 1) * 4);
    *Eb = PyByteArray_FromStringAndSize(NULL, (Py_ssize_t)(*total > 0 ? *total : 1) * 4);
    if (!*Vb || !*Eb) return 2;
    j->E = (int32_t*)PyByteArray_AS_STRING(*Eb);
    j->cut[0] = 0;
    Py_ssize_t r = 0;
    for (int t = 0; t < j->T; ++t) {
        /* up to the first row whose arcs start at or past (t + 1) / T of them */
        const int64_t goal = *total * (t + 1) / j->T;
        if (t == j->T - 1) r = n;
        else
            while (r < n && j->off[r] < goal) ++r;
        j->cut[t + 1] = r;
    }
    return 0;
}

static PyObject* int_arcs(PyObject* self, PyObject* args) {
    (void)self;
    PyObject* adj;
    Py_ssize_t n;
    int drop_self, threads;
    if (!PyArg_ParseTuple(args, "Onpi", &adj, &n, &drop_self, &threads)) return NULL;
    if (!PyDict_Check(adj) || n < 0 || PyDict_GET_SIZE(adj) != n) Py_RETURN_NONE;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    PyObject** keys = (PyObject**)malloc(sizeof(PyObject*) * nn);
    PyObject** rows = (PyObject**)malloc(sizeof(PyObject*) * nn);
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t* kept = (int32_t*)malloc(sizeof(int32_t) * nn);
    PyObject *Vb = NULL, *Eb = NULL, *res = NULL;
    IntArcsJob job;
    IntArcsPart parts[INT_ARCS_MAX_THREADS];
    memset(&job, 0, sizeof(job));
    if (!keys || !rows || !off || !kept) {
        PyErr_NoMemory();
        goto out;
    }
    /* the adjacency's entries in order; the parts check them */
    Py_ssize_t pos = 0, i = 0;
    PyObject *u, *nbrs;
    while (i < n && PyDict_Next(adj, &pos, &u, &nbrs)) {
        keys[i] = u;
        rows[i] = nbrs;
        ++i;
    }
    if (i != n) {
        res = Py_None;
        goto out;
    }
    int T = threads < 1 ? 1 : (threads > INT_ARCS_MAX_THREADS ? INT_ARCS_MAX_THREADS : threads);
    if (n < 4096) T = 1; /* few rows: not worth the threads */
    if (T > 1) T = 1 + pool_ready(T);
    job.keys = keys;
    job.rows = rows;
    job.off = off;
    job.kept = kept;
    job.n = n;
    job.drop_self = drop_self;
    job.T = T;
    for (int t = 0; t < T; ++t) parts[t] = (IntArcsPart){&job, t, 0, 0, 0};
    /* the calling thread keeps the GIL through both phases */
    run_parallel(T, int_arcs_sizes, parts, sizeof(IntArcsPart));
    int64_t total = 0;
    int bad = int_arcs_setup(&job, parts, &Vb, &Eb, &total);
    if (bad == 2) goto out; /* the allocation's exception is set */
    if (bad) {
        res = Py_None;
        goto out;
    }
    run_parallel(T, int_arcs_fill, parts, sizeof(IntArcsPart));
    uint64_t sum = 0, xr = 0x243F6A8885A308D3ull;
    for (int t = 0; t < T; ++t) {
        bad |= parts[t].bad;
        sum += parts[t].sum;
        xr ^= parts[t].xr;
    }
    if (bad) {
        res = Py_None;
        goto out;
    }
    /* rows that dropped a self-loop leave gaps: close them, in order */
    int32_t* V = (int32_t*)PyByteArray_AS_STRING(Vb);
    int32_t* E = job.E;
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
    free(keys);
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
