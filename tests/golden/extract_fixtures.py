"""Extract the reference's simulation regression fixtures into plain .npz files.

The reference stores its golden outputs as gzip'd pickles
(tests/expected/sim/{rWW,JR,Kuramoto,WC}.pkl.gz, written by
tests/sim/gen_expected.py:19-57).  Unpickling untrusted data executes code, so
this script never calls ``pickle.load``: it walks the opcode stream with
``pickletools.genops`` (a disassembler that executes nothing) and evaluates a
tiny symbolic stack machine that only understands the opcodes those files use.
NumPy arrays appear as ``_reconstruct(ndarray, (0,), b'b')`` followed by a
BUILD whose state tuple is ``(version, shape, dtype, fortran, raw_bytes)``;
the raw bytes are decoded with ``np.frombuffer``.

The test inputs the fixtures were produced from are data files bundled with
the reference (HCP group-train706 Schaefer-100 SC strength and length,
src/cubnm/data/hcp/sc/group-train706/*.npz, loaded with allow_pickle=False).

Usage (in a container that has the reference mounted):
    python tests/golden/extract_fixtures.py /root/reference
Writes tests/golden/sim_<Model>.npz and tests/golden/hcp_sc_*.npz.
"""
import gzip
import os
import pickletools
import sys

import numpy as np


class _Global:
    def __init__(self, module, name):
        self.module, self.name = module, name

    def __repr__(self):
        return f"{self.module}.{self.name}"


class _Call:
    def __init__(self, fn, args):
        self.fn, self.args, self.state = fn, args, None


_MARK = object()


def _materialise(obj):
    """Turn symbolic calls into numpy arrays / dtypes, recursively."""
    if isinstance(obj, dict):
        return {_materialise(k): _materialise(v) for k, v in obj.items()}
    if isinstance(obj, tuple):
        return tuple(_materialise(o) for o in obj)
    if isinstance(obj, list):
        return [_materialise(o) for o in obj]
    if isinstance(obj, _Call):
        fn = obj.fn
        if isinstance(fn, _Global) and fn.name == "dtype":
            return np.dtype(obj.args[0])
        if isinstance(fn, _Global) and fn.name == "_reconstruct":
            _, shape, dtype, fortran, raw = obj.state
            dtype = _materialise(dtype)
            if dtype.byteorder == "=" or dtype.byteorder == "|":
                pass
            arr = np.frombuffer(raw, dtype=dtype.newbyteorder("<")).copy()
            order = "F" if fortran else "C"
            return arr.reshape(shape, order=order)
        if isinstance(fn, _Global) and fn.name == "scalar":
            dt = _materialise(obj.args[0])
            return np.frombuffer(obj.args[1], dtype=dt)[0]
        raise ValueError(f"unsupported call {fn!r}")
    return obj


def decode(data):
    stack, memo = [], {}
    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            break
        if name == "EMPTY_DICT":
            stack.append({})
        elif name == "EMPTY_LIST":
            stack.append([])
        elif name == "EMPTY_TUPLE":
            stack.append(())
        elif name == "MARK":
            stack.append(_MARK)
        elif name in ("SHORT_BINUNICODE", "BINUNICODE", "SHORT_BINBYTES", "BINBYTES",
                      "BININT1", "BININT2", "BININT", "BINFLOAT", "LONG1"):
            stack.append(arg)
        elif name == "NONE":
            stack.append(None)
        elif name == "NEWTRUE":
            stack.append(True)
        elif name == "NEWFALSE":
            stack.append(False)
        elif name == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif name in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif name in ("BINPUT", "LONG_BINPUT"):
            memo[arg] = stack[-1]
        elif name == "STACK_GLOBAL":
            nm = stack.pop()
            mod = stack.pop()
            stack.append(_Global(mod, nm))
        elif name == "GLOBAL":
            mod, nm = arg.split(" ")
            stack.append(_Global(mod, nm))
        elif name in ("TUPLE1", "TUPLE2", "TUPLE3"):
            n = int(name[-1])
            items = tuple(stack[-n:])
            del stack[-n:]
            stack.append(items)
        elif name == "TUPLE":
            i = len(stack) - 1 - stack[::-1].index(_MARK)
            items = tuple(stack[i + 1:])
            del stack[i:]
            stack.append(items)
        elif name == "REDUCE":
            args = stack.pop()
            fn = stack.pop()
            stack.append(_Call(fn, args))
        elif name == "BUILD":
            state = stack.pop()
            obj = stack[-1]
            if not isinstance(obj, _Call):
                raise ValueError("BUILD on a non-call object")
            obj.state = state
        elif name == "SETITEMS":
            i = len(stack) - 1 - stack[::-1].index(_MARK)
            items = stack[i + 1:]
            del stack[i:]
            d = stack[-1]
            for k, v in zip(items[0::2], items[1::2]):
                d[k] = v
        elif name == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif name == "APPENDS":
            i = len(stack) - 1 - stack[::-1].index(_MARK)
            items = stack[i + 1:]
            del stack[i:]
            stack[-1].extend(items)
        else:
            raise ValueError(f"opcode {name} not supported by this decoder")
    assert len(stack) == 1
    return _materialise(stack[0])


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for model in ["rWW", "JR", "Kuramoto", "WC", "rWWEx"]:
        path = os.path.join(ref_root, "tests", "expected", "sim", f"{model}.pkl.gz")
        with gzip.open(path, "rb") as f:
            data = f.read()
        fixtures = decode(data)
        flat = {}
        for key, entry in fixtures.items():
            for field, arr in entry.items():
                flat[f"{key}|{field}"] = np.asarray(arr, dtype=np.float64)
        np.savez_compressed(os.path.join(here, f"sim_{model}.npz"), **flat)
        print(model, len(fixtures), "entries")
    sc_dir = os.path.join(ref_root, "src", "cubnm", "data", "hcp", "sc", "group-train706")
    for measure in ["strength", "length"]:
        mat = np.load(os.path.join(sc_dir, f"ctx_parc-schaefer-100_desc-{measure}.npz"),
                      allow_pickle=False)["arr_0"]
        np.savez_compressed(os.path.join(here, f"hcp_sc_{measure}.npz"), mat=np.asarray(mat, np.float64))
        print(measure, mat.shape, mat.dtype)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
