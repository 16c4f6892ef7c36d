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
 * ints_from_rows(buffer) -> list of the 256-bit little-endian values of 32-byte rows.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

/* 1 if v is an exact int in [0, 2^256), written to out; 0 if it needs the interpreter. */
static int fast_row(PyObject* v, unsigned char* out) {
  if (!PyLong_CheckExact(v) || _PyLong_Sign(v) < 0) return 0;
  if (_PyLong_NumBits(v) > 256) return 0;
  return _PyLong_AsByteArray((PyLongObject*)v, out, 32, 1, 0) == 0;
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
  unsigned char* buf;
  unsigned char* done; /* per element: 1 when a worker wrote its row */
  Py_ssize_t lo, hi;
} Span;

static void* convert_span(void* arg) {
  Span* s = (Span*)arg;
  for (Py_ssize_t i = s->lo; i < s->hi; ++i) s->done[i] = (unsigned char)fast_row(s->items[i], s->buf + 32 * i);
  return NULL;
}

enum { kMaxThreads = 16, kMinPerThread = 1 << 15 };

static int ingest_threads(Py_ssize_t n) {
  long cpus = sysconf(_SC_NPROCESSORS_ONLN);
  const char* env = getenv("ZK_INGEST_THREADS");
  if (env && atoi(env) > 0) cpus = atoi(env);
  if (cpus > kMaxThreads) cpus = kMaxThreads;
  Py_ssize_t t = n / kMinPerThread;
  if (t > cpus) t = cpus;
  return t < 1 ? 1 : (int)t;
}

static PyObject* rows_from_ints(PyObject* self, PyObject* args) {
  (void)self;
  PyObject *values, *modulus;
  const char* name;
  if (!PyArg_ParseTuple(args, "OOs", &values, &modulus, &name)) return NULL;
  if (!PyLong_Check(modulus) || _PyLong_Sign(modulus) <= 0) {
    PyErr_SetString(PyExc_ValueError, "modulus must be a positive int");
    return NULL;
  }
  PyObject* items = PySequence_Fast(values, "values must be iterable");
  if (!items) return NULL;
  const Py_ssize_t n = PySequence_Fast_GET_SIZE(items);
  PyObject* out = PyBytes_FromStringAndSize(NULL, 32 * n);
  unsigned char* done = n ? (unsigned char*)calloc((size_t)n, 1) : NULL;
  if (!out || (n && !done)) {
    Py_XDECREF(out);
    Py_DECREF(items);
    free(done);
    return PyErr_NoMemory();
  }
  unsigned char* buf = (unsigned char*)PyBytes_AS_STRING(out);
  PyObject** it = PySequence_Fast_ITEMS(items);
  const int nt = ingest_threads(n);
  Span spans[kMaxThreads];
  pthread_t tids[kMaxThreads];
  int started = 0;
  for (int t = 0; t < nt; ++t) {
    spans[t] = (Span){it, buf, done, n * t / nt, n * (t + 1) / nt};
    if (t > 0 && pthread_create(&tids[t], NULL, convert_span, &spans[t]) == 0) ++started;
    else if (t > 0) break; /* fewer workers: the calling thread's second pass converts the rest */
  }
  convert_span(&spans[0]);
  for (int t = 1; t <= started; ++t) pthread_join(tids[t], NULL);
  for (Py_ssize_t i = 0; i < n; ++i) {
    if (done[i]) continue;
    if (slow_row(it[i], modulus, name, buf + 32 * i) < 0) {
      free(done);
      Py_DECREF(out);
      Py_DECREF(items);
      return NULL;
    }
  }
  free(done);
  PyObject* res = PyTuple_Pack(2, out, items);
  Py_DECREF(out);
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
    {"ints_from_rows", ints_from_rows, METH_O, "buffer of 32-byte rows -> list of ints"},
    {NULL, NULL, 0, NULL},
};

static struct PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_ingest", NULL, -1, kMethods};

PyMODINIT_FUNC PyInit__ingest(void) { return PyModule_Create(&kModule); }
