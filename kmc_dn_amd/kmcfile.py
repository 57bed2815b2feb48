"""Reader/writer for ``.kmc`` state files (kmc_dn.saveSelf / loadSelf format).

The reference writes a pickled dict of every list/tuple/int/float/ndarray
attribute (kmc_dopant_networks.py:943-963) and reads it back with
``pickle.load`` (:965-978).  Loading a pickle runs whatever the file names, so
this module never unpickles: :func:`read_kmc` walks the opcode stream with
``pickletools.genops`` (a pure disassembler) on a small stack machine that
understands only literals, containers, memo references and the four numpy
reconstruction patterns ``saveSelf`` produces (ndarray via
``numpy.core.multiarray._reconstruct`` + BUILD, ``numpy.dtype`` + BUILD,
``numpy.core.multiarray.scalar``).  Arrays are rebuilt from their raw bytes with
``np.frombuffer``; no global is ever imported or called, and anything else in
the stream (object arrays, other globals) raises :class:`KmcFormatError`.

:func:`write_kmc` emits the same layout (protocol 3, numpy arrays pickled the
standard way) so files round-trip with the reference's ``loadSelf``.
"""
from __future__ import annotations

import io
import pickle
import pickletools
from typing import Any

import numpy as np


class KmcFormatError(ValueError):
    pass


_MARK = object()


class _Global:
    __slots__ = ("module", "name")

    def __init__(self, module: str, name: str):
        self.module, self.name = module, name

    def key(self):
        mod = self.module.replace("numpy._core", "numpy.core")
        return mod, self.name


class _Call:
    """A REDUCE of an allowed numpy global, resolved after BUILD."""
    __slots__ = ("func", "args", "state", "value", "done")

    def __init__(self, func: _Global, args: tuple):
        self.func, self.args, self.state, self.value, self.done = func, args, None, None, False


_ALLOWED = {("numpy.core.multiarray", "_reconstruct"), ("numpy", "ndarray"),
            ("numpy", "dtype"), ("numpy.core.multiarray", "scalar")}


def _resolve(obj: Any) -> Any:
    if isinstance(obj, _Call):
        if not obj.done:
            obj.value = _materialise(obj)
            obj.done = True
        return obj.value
    if isinstance(obj, dict):
        return {_resolve(k): _resolve(v) for k, v in obj.items()}
    if isinstance(obj, list):
        return [_resolve(v) for v in obj]
    if isinstance(obj, tuple):
        return tuple(_resolve(v) for v in obj)
    if isinstance(obj, _Global):
        raise KmcFormatError(f"bare global {obj.module}.{obj.name} in stream")
    return obj


def _dtype_of(call: _Call) -> np.dtype:
    if call.func.key() != ("numpy", "dtype"):
        raise KmcFormatError("expected numpy.dtype")
    name = call.args[0]
    if not isinstance(name, str) or name.startswith("O") or name.startswith("V"):
        raise KmcFormatError(f"unsupported dtype {name!r}")
    dt = np.dtype(name)
    st = call.state
    if st is not None:
        order = st[1]
        if order in ("<", ">"):
            dt = dt.newbyteorder(order)
        if any(x is not None for x in st[2:5]):
            raise KmcFormatError("structured dtypes are not supported")
    return dt


def _materialise(call: _Call) -> Any:
    key = call.func.key()
    if key == ("numpy", "dtype"):
        return _dtype_of(call)
    if key == ("numpy.core.multiarray", "scalar"):
        dt_call, raw = call.args
        if not isinstance(dt_call, _Call) or not isinstance(raw, (bytes, bytearray)):
            raise KmcFormatError("bad numpy scalar")
        return np.frombuffer(bytes(raw), dtype=_dtype_of(dt_call))[0]
    if key == ("numpy.core.multiarray", "_reconstruct"):
        cls = call.args[0]
        if not isinstance(cls, _Global) or cls.key() != ("numpy", "ndarray"):
            raise KmcFormatError("only plain ndarrays are supported")
        st = call.state
        if st is None or len(st) != 5:
            raise KmcFormatError("ndarray without state")
        _, shape, dt_call, fortran, raw = st
        if not isinstance(dt_call, _Call):
            raise KmcFormatError("ndarray dtype is not a numpy.dtype")
        dt = _dtype_of(dt_call)
        if not isinstance(raw, (bytes, bytearray)):
            raise KmcFormatError("object arrays are not supported")
        shape = tuple(int(s) for s in shape)
        arr = np.frombuffer(bytes(raw), dtype=dt)
        arr = arr.reshape(shape, order="F" if fortran else "C")
        return arr.astype(dt.newbyteorder("="), copy=True)
    raise KmcFormatError(f"global {key} not allowed")


def _decode(data: bytes) -> Any:
    stack: list = []
    memo: dict = {}

    def pop_mark():
        items = []
        while True:
            x = stack.pop()
            if x is _MARK:
                break
            items.append(x)
        items.reverse()
        return items

    for op, arg, _pos in pickletools.genops(data):
        n = op.name
        if n in ("PROTO", "FRAME"):
            continue
        elif n == "STOP":
            break
        elif n == "MARK":
            stack.append(_MARK)
        elif n in ("EMPTY_DICT",):
            stack.append({})
        elif n == "EMPTY_LIST":
            stack.append([])
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n in ("BINUNICODE", "SHORT_BINUNICODE", "BINUNICODE8", "UNICODE"):
            stack.append(str(arg))
        elif n in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG"):
            stack.append(int(arg))
        elif n in ("BINFLOAT", "FLOAT"):
            stack.append(float(arg))
        elif n in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8"):
            stack.append(bytes(arg))
        elif n == "NONE":
            stack.append(None)
        elif n == "NEWTRUE":
            stack.append(True)
        elif n == "NEWFALSE":
            stack.append(False)
        elif n == "TUPLE1":
            stack.append((stack.pop(),))
        elif n == "TUPLE2":
            b = stack.pop(); a = stack.pop(); stack.append((a, b))
        elif n == "TUPLE3":
            c = stack.pop(); b = stack.pop(); a = stack.pop(); stack.append((a, b, c))
        elif n == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif n == "LIST":
            stack.append(list(pop_mark()))
        elif n == "DICT":
            items = pop_mark()
            stack.append({items[i]: items[i + 1] for i in range(0, len(items), 2)})
        elif n == "APPEND":
            v = stack.pop(); stack[-1].append(v)
        elif n == "APPENDS":
            items = pop_mark(); stack[-1].extend(items)
        elif n == "SETITEM":
            v = stack.pop(); k = stack.pop(); stack[-1][k] = v
        elif n == "SETITEMS":
            items = pop_mark()
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[items[i]] = items[i + 1]
        elif n in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif n == "GLOBAL":
            module, name = arg.split(" ", 1) if " " in arg else arg.split("\n", 1)
            g = _Global(module, name)
            if g.key() not in _ALLOWED:
                raise KmcFormatError(f"global {module}.{name} not allowed")
            stack.append(g)
        elif n == "STACK_GLOBAL":
            name = stack.pop(); module = stack.pop()
            g = _Global(module, name)
            if g.key() not in _ALLOWED:
                raise KmcFormatError(f"global {module}.{name} not allowed")
            stack.append(g)
        elif n == "REDUCE":
            args = stack.pop(); func = stack.pop()
            if not isinstance(func, _Global):
                raise KmcFormatError("REDUCE of a non-global")
            stack.append(_Call(func, args))
        elif n == "BUILD":
            state = stack.pop()
            target = stack[-1]
            if not isinstance(target, _Call):
                raise KmcFormatError("BUILD on a non-numpy object")
            target.state = state
        else:
            raise KmcFormatError(f"opcode {n} not supported")
    if len(stack) != 1:
        raise KmcFormatError("malformed stream")
    return _resolve(stack[0])


def read_kmc(path_or_bytes) -> dict:
    """Decode a ``.kmc`` file into a dict of attributes (numpy arrays, numbers, lists)."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            data = f.read()
    out = _decode(data)
    if not isinstance(out, dict):
        raise KmcFormatError("top-level object is not a dict")
    return out


def write_kmc(path, attrs: dict) -> None:
    """Write ``attrs`` in the saveSelf layout (only list/tuple/int/float/ndarray values)."""
    d = {}
    for k, v in attrs.items():
        if isinstance(v, np.ndarray) and v.dtype == object:
            continue
        if isinstance(v, (list, tuple, int, float, np.ndarray)):
            d[k] = v
    buf = io.BytesIO()
    pickle.dump(d, buf, protocol=3)
    with open(path, "wb") as f:
        f.write(buf.getvalue())
