/*
 * zkgpu._ingest: bulk conversion between Python ints and the (n, 4) little-endian uint64 rows the
 * device kernels read, for the public API's int inputs and outputs (Field256.from_ints / to_ints,
 * msm_pippenger_gpu). The reference converts one element at a time in Python
 * (src/zkgpu/field256_gpu.py:239-247, msm_gpu.py:432-442); here one C loop does it.
 *
 * rows_from_ints(values, modulus, name) -> (bytes, items)
 *   values: any iterable of integers (int, bool, or objects with __index__, as operator.index
 *   accepts). A value in [0, 2^256) is copied as it is (the device reduces it); a negative value
 *   or one of 2^256 and more is reduced mod `modulus` here. A non-integer raises
 *   TypeError("<name> must contain integers: <operator.index's message>"), the reference's
 *   message (operator.index's own TypeError when name is empty). items is the sequence the
 *   values were read from (a list when values was not one), for messages that quote an element.
 * ints_from_rows(buffer) -> list of the 256-bit little-endian values of 32-byte rows.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>

static int row_of(PyObject* v, PyObject* modulus, unsigned char* out) {
  const int neg = _PyLong_Sign(v) < 0;
  const size_t bits = neg ? 0 : _PyLong_NumBits(v);
  if (bits == (size_t)-1 && PyErr_Occurred()) return -1;
  if (!neg && bits <= 256) return _PyLong_AsByteArray((PyLongObject*)v, out, 32, 1, 0);
  PyObject* r = PyNumber_Remainder(v, modulus);
  if (!r) return -1;
  const int rc = _PyLong_AsByteArray((PyLongObject*)r, out, 32, 1, 0);
  Py_DECREF(r);
  return rc;
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
  if (!out) {
    Py_DECREF(items);
    return NULL;
  }
  unsigned char* buf = (unsigned char*)PyBytes_AS_STRING(out);
  PyObject** it = PySequence_Fast_ITEMS(items);
  for (Py_ssize_t i = 0; i < n; ++i) {
    PyObject* v = PyLong_CheckExact(it[i]) ? (Py_INCREF(it[i]), it[i]) : PyNumber_Index(it[i]);
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
      Py_DECREF(out);
      Py_DECREF(items);
      return NULL;
    }
    const int rc = row_of(v, modulus, buf + 32 * i);
    Py_DECREF(v);
    if (rc < 0) {
      Py_DECREF(out);
      Py_DECREF(items);
      return NULL;
    }
  }
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
