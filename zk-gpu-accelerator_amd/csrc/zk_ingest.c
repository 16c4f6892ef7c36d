/*
 * zkgpu._ingest: bulk conversion between Python ints and the (n, 4) little-endian uint64 rows the
 * device kernels read, for the public API's int inputs and outputs (Field256.from_ints / to_ints,
 * msm_pippenger_gpu). The reference converts one element at a time in Python
 * (src/zkgpu/field256_gpu.py:239-247, msm_gpu.py:432-442); here C loops do it.
 *
 * rows_from_ints(values, modulus, name) -> (bytes, items)
 *   values: any iterable of integers (int, bool, or objects with __index__, as operator.index
 *   accepts). A value in [0, 2^256) is copied as it is (the device reduces it); a negative value
 *   or one of 2^256 and more is reduced mod `modulus` here. A non-integer raises
 *   TypeError("<name> must contain integers: <operator.index's message>"), the reference's
 *   message (operator.index's own TypeError when name is empty). items is the sequence the
 *   values were read from (a list when values was not one), for messages that quote an element.
 *   Long inputs are split over worker threads. The caller holds the interpreter lock for the
 *   whole call, so no other Python code runs and every element stays alive and unchanged; the
 *   workers only read exact ints in [0, 2^256) (type, sign, bit length and digits, none of which
 *   allocates or raises) and leave every other element to the calling thread, which converts
 *   those afterwards in index order with the interpreter's own functions.
 * rows_into(values, modulus, name, dest, stride, offset[, start, stop]) -> (count, items)
 *   The same conversion written into a writable buffer (page-locked staging memory for an upload):
 *   row i goes to 32-byte row offset + i * stride of dest. With start and stop, only the elements
 *   [start, stop) of the sequence (clamped to its length) are converted, so that a caller can
 *   upload one part while the next is converted; count is always the sequence's length.
 *   ValueError if dest is too small.
 * ints_from_rows(buffer) -> list of the 256-bit little-endian values of 32-byte rows.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* CPython before 3.12 keeps an int as Py_SIZE(v) 30-bit digits (the sign in the size): reading them
 * directly packs a row in a few shifts instead of a general byte-array conversion. */
#if PY_VERSION_HEX < 0x030C0000 && PYLONG_BITS_IN_DIGIT == 30 && PY_LITTLE_ENDIAN
#include <longintrepr.h>
#define ZK_INGEST_DIGITS 1
#endif

/* 1 if v is an exact int in [0, 2^256), written to out; 0 if it needs the interpreter. */
static int fast_row(PyObject* v, unsigned char* out) {
  if (!PyLong_CheckExact(v)) return 0;
#ifdef ZK_INGEST_DIGITS
  const Py_ssize_t size = Py_SIZE(v);
  if (size < 0 || size > 9) return 0; /* negative, or at least 2^270 */
  const digit* d = ((PyLongObject*)v)->ob_digit;
  if (size == 9 && (d[8] >> 16)) return 0; /* at least 2^256 */
  uint64_t w[5] = {0, 0, 0, 0, 0};
  for (Py_ssize_t i = 0; i < size; ++i) {
    const unsigned bit = 30u * (unsigned)i, q = bit >> 6, r = bit & 63u;
    w[q] |= (uint64_t)d[i] << r;
    if (r > 34) w[q + 1] |= (uint64_t)d[i] >> (64u - r);
  }
  memcpy(out, w, 32);
  return 1;
#else
  if (_PyLong_Sign(v) < 0) return 0;
  if (_PyLong_NumBits(v) > 256) return 0;
  return _PyLong_AsByteArray((PyLongObject*)v, out, 32, 1, 0) == 0;
#endif
}

static int slow_row(PyObject* item, PyObject* modulus, const char* name, unsigned char* out) {
  PyObject* v = PyLong_CheckExact(item) ? (Py_INCREF(item), item) : PyNumber_Index(item);
  if (!v) {
    if (name[0] && PyErr_ExceptionMatches(PyExc_TypeError)) {
      PyObject *type, *value, *tb;
      PyErr_Fetch(&type, &value, &tb);
      PyErr_NormalizeException(&type, &value, &tb);
      PyErr_Format(PyExc_TypeError, "%s must contain integers: %S", name, value);
      Py_XDECREF(type);
      Py_XDECREF(value);
      Py_XDECREF(tb);
    }
    return -1;
  }
  int rc;
  const int neg = _PyLong_Sign(v) < 0;
  const size_t bits = neg ? 0 : _PyLong_NumBits(v);
  if (bits == (size_t)-1 && PyErr_Occurred()) {
    rc = -1;
  } else if (!neg && bits <= 256) {
    rc = _PyLong_AsByteArray((PyLongObject*)v, out, 32, 1, 0);
  } else {
    PyObject* r = PyNumber_Remainder(v, modulus);
    rc = r ? _PyLong_AsByteArray((PyLongObject*)r, out, 32, 1, 0) : -1;
    Py_XDECREF(r);
  }
  Py_DECREF(v);
  return rc;
}

typedef struct {
  PyObject** items;
  unsigned char* buf;  /* row i at buf + 32 * stride * i */
  Py_ssize_t stride;
  unsigned char* done; /* per element: 1 when a worker wrote its row */
  Py_ssize_t lo, hi;
} Span;

static void* convert_span(void* arg) {
  Span* s = (Span*)arg;
  const Py_ssize_t step = 32 * s->stride;
  for (Py_ssize_t i = s->lo; i < s->hi; ++i) s->done[i] = (unsigned char)fast_row(s->items[i], s->buf + step * i);
  return NULL;
}

enum { kMaxThreads = 16, kMinPerThread = 1 << 13 };

static int ingest_threads(Py_ssize_t n) {
  long cpus = sysconf(_SC_NPROCESSORS_ONLN);
  const char* env = getenv("ZK_INGEST_THREADS");
  if (env && atoi(env) > 0) cpus = atoi(env);
  if (cpus > kMaxThreads) cpus = kMaxThreads;
  Py_ssize_t t = n / kMinPerThread;
  if (t > cpus) t = cpus;
  return t < 1 ? 1 : (int)t;
}

/* Worker threads are started once and kept, so that a caller converting in parts (one part
 * uploaded while the next is converted) pays a wake-up per part rather than thread creations.
 * A job is published under the lock with a new generation number; every started worker runs its
 * span (an empty one past the job's span count) and reports back. The caller holds the
 * interpreter lock throughout, so jobs never overlap. A forked child has none of the parent's
 * workers: the pool starts empty there. */
static struct {
  pthread_mutex_t mu;
  pthread_cond_t go, done;
  int workers;         /* started, ids 1..workers */
  unsigned long gen;   /* the current job */
  int pending;         /* workers still on the current job */
  int nspans;
  Span spans[kMaxThreads];
  unsigned long born[kMaxThreads]; /* the generation before a worker's first job */
} g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, 0, 0, {{0}}, {0}};

static void* pool_worker(void* arg) {
  const int id = (int)(intptr_t)arg;
  pthread_mutex_lock(&g_pool.mu);
  unsigned long seen = g_pool.born[id]; /* the job published right after its creation is new to it */
  for (;;) {
    while (g_pool.gen == seen) pthread_cond_wait(&g_pool.go, &g_pool.mu);
    seen = g_pool.gen;
    Span span = g_pool.spans[id];
    const int active = id < g_pool.nspans;
    pthread_mutex_unlock(&g_pool.mu);
    if (active) convert_span(&span);
    pthread_mutex_lock(&g_pool.mu);
    if (--g_pool.pending == 0) pthread_cond_signal(&g_pool.done);
  }
  return NULL;
}

static void pool_after_fork_in_child(void) {
  pthread_mutex_init(&g_pool.mu, NULL);
  pthread_cond_init(&g_pool.go, NULL);
  pthread_cond_init(&g_pool.done, NULL);
  g_pool.workers = 0;
  g_pool.pending = 0;
}

/* Runs spans[0, nt) with the calling thread taking span 0; returns once every span is done. */
static void pool_run(Span* spans, int nt) {
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.workers < nt - 1) {
    /* a new worker may first take the lock after this job is published: it starts from the
       generation before it, so the job counts as new to it */
    pthread_t tid;
    g_pool.born[g_pool.workers + 1] = g_pool.gen;
    if (pthread_create(&tid, NULL, pool_worker, (void*)(intptr_t)(g_pool.workers + 1)) != 0) break;
    pthread_detach(tid);
    ++g_pool.workers;
  }
  const int nw = g_pool.workers;
  const int run = nt - 1 < nw ? nt - 1 : nw; /* spans the workers take; the rest stay undone */
  for (int t = 1; t <= run; ++t) g_pool.spans[t] = spans[t];
  g_pool.nspans = run + 1;
  g_pool.pending = nw;
  ++g_pool.gen;
  pthread_cond_broadcast(&g_pool.go);
  pthread_mutex_unlock(&g_pool.mu);
  convert_span(&spans[0]);
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.pending > 0) pthread_cond_wait(&g_pool.done, &g_pool.mu);
  pthread_mutex_unlock(&g_pool.mu);
}

/* Converts the n items into rows buf + 32 * stride * i; 0, or -1 with the exception set. */
static int convert_items(PyObject** it, Py_ssize_t n, PyObject* modulus, const char* name, unsigned char* buf,
                         Py_ssize_t stride) {
  unsigned char* done = n ? (unsigned char*)calloc((size_t)n, 1) : NULL;
  if (n && !done) {
    PyErr_NoMemory();
    return -1;
  }
  const int nt = ingest_threads(n);
  Span spans[kMaxThreads];
  for (int t = 0; t < nt; ++t) spans[t] = (Span){it, buf, stride, done, n * t / nt, n * (t + 1) / nt};
  if (nt > 1) pool_run(spans, nt);
  else convert_span(&spans[0]);
  /* what no worker converted (other objects, values to reduce, spans without a worker) */
  for (Py_ssize_t i = 0; i < n; ++i) {
    if (done[i]) continue;
    if (slow_row(it[i], modulus, name, buf + 32 * stride * i) < 0) {
      free(done);
      return -1;
    }
  }
  free(done);
  return 0;
}

static int check_modulus(PyObject* modulus) {
  if (!PyLong_Check(modulus) || _PyLong_Sign(modulus) <= 0) {
    PyErr_SetString(PyExc_ValueError, "modulus must be a positive int");
    return -1;
  }
  return 0;
}

static PyObject* rows_from_ints(PyObject* self, PyObject* args) {
  (void)self;
  PyObject *values, *modulus;
  const char* name;
  if (!PyArg_ParseTuple(args, "OOs", &values, &modulus, &name)) return NULL;
  if (check_modulus(modulus) < 0) return NULL;
  PyObject* items = PySequence_Fast(values, "values must be iterable");
  if (!items) return NULL;
  const Py_ssize_t n = PySequence_Fast_GET_SIZE(items);
  PyObject* out = PyBytes_FromStringAndSize(NULL, 32 * n);
  if (!out) {
    Py_DECREF(items);
    return NULL;
  }
  if (convert_items(PySequence_Fast_ITEMS(items), n, modulus, name, (unsigned char*)PyBytes_AS_STRING(out), 1) < 0) {
    Py_DECREF(out);
    Py_DECREF(items);
    return NULL;
  }
  PyObject* res = PyTuple_Pack(2, out, items);
  Py_DECREF(out);
  Py_DECREF(items);
  return res;
}

static PyObject* rows_into(PyObject* self, PyObject* args) {
  (void)self;
  PyObject *values, *modulus, *dest;
  const char* name;
  Py_ssize_t stride, offset, start = 0, stop = PY_SSIZE_T_MAX;
  if (!PyArg_ParseTuple(args, "OOsOnn|nn", &values, &modulus, &name, &dest, &stride, &offset, &start, &stop)) return NULL;
  if (check_modulus(modulus) < 0) return NULL;
  if (stride < 1 || offset < 0 || start < 0) {
    PyErr_SetString(PyExc_ValueError, "stride must be positive, offset and start non-negative");
    return NULL;
  }
  PyObject* items = PySequence_Fast(values, "values must be iterable");
  if (!items) return NULL;
  const Py_ssize_t n = PySequence_Fast_GET_SIZE(items);
  if (stop > n) stop = n;
  if (start > stop) start = stop;
  Py_buffer view;
  if (PyObject_GetBuffer(dest, &view, PyBUF_WRITABLE | PyBUF_C_CONTIGUOUS) < 0) {
    Py_DECREF(items);
    return NULL;
  }
  const Py_ssize_t rows = view.len / 32;
  if (n && offset + (n - 1) * stride >= rows) {
    PyBuffer_Release(&view);
    Py_DECREF(items);
    PyErr_SetString(PyExc_ValueError, "rows_into: the destination is too small");
    return NULL;
  }
  const int rc = convert_items(PySequence_Fast_ITEMS(items) + start, stop - start, modulus, name,
                               (unsigned char*)view.buf + 32 * (offset + start * stride), stride);
  PyBuffer_Release(&view);
  if (rc < 0) {
    Py_DECREF(items);
    return NULL;
  }
  PyObject* res = Py_BuildValue("(nO)", n, items);
  Py_DECREF(items);
  return res;
}

static PyObject* ints_from_rows(PyObject* self, PyObject* arg) {
  (void)self;
  Py_buffer view;
  if (PyObject_GetBuffer(arg, &view, PyBUF_C_CONTIGUOUS) < 0) return NULL;
  if (view.len % 32) {
    PyBuffer_Release(&view);
    PyErr_SetString(PyExc_ValueError, "rows must be a whole number of 32-byte rows");
    return NULL;
  }
  const Py_ssize_t n = view.len / 32;
  PyObject* list = PyList_New(n);
  if (!list) {
    PyBuffer_Release(&view);
    return NULL;
  }
  const unsigned char* p = (const unsigned char*)view.buf;
  for (Py_ssize_t i = 0; i < n; ++i) {
    PyObject* v = _PyLong_FromByteArray(p + 32 * i, 32, 1, 0);
    if (!v) {
      Py_DECREF(list);
      PyBuffer_Release(&view);
      return NULL;
    }
    PyList_SET_ITEM(list, i, v);
  }
  PyBuffer_Release(&view);
  return list;
}

static PyMethodDef kMethods[] = {
    {"rows_from_ints", rows_from_ints, METH_VARARGS, "(values, modulus, name) -> (32-byte rows as bytes, items)"},
    {"rows_into", rows_into, METH_VARARGS, "(values, modulus, name, dest, stride, offset) -> (count, items)"},
    {"ints_from_rows", ints_from_rows, METH_O, "buffer of 32-byte rows -> list of ints"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_ingest", NULL, -1, kMethods};

PyMODINIT_FUNC PyInit__ingest(void) {
  static int registered = 0;
  if (!registered && pthread_atfork(NULL, NULL, pool_after_fork_in_child) == 0) registered = 1;
  return PyModule_Create(&kModule);
}
