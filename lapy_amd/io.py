"""ShapeDNA text formats: ``.ev`` spectra and PSOL vertex functions.

Host-side file IO for the eigensolver's results (reference: ``lapy/io.py``).
The functions keep the reference's names, arguments, dictionary keys and
error types, and write byte-for-byte the text the reference writes, so that
the GPU results drop into existing ShapeDNA/BrainPrint tooling:

- ``read_vfunc`` / ``write_vfunc``  (``lapy/io.py:10-53``, ``:287-310``)
- ``read_ev`` / ``write_ev``        (``lapy/io.py:56-186``, ``:189-284``)

Nothing here touches the GPU: the eigenpairs arrive as numpy arrays from
``Solver.eigs`` or ``shapedna.compute_shapedna``.
"""

import logging
import re

import numpy as np

logger = logging.getLogger(__name__)

# Header fields of an .ev file, in the order they are written, and the cast
# applied when they are read back (lapy/io.py:96-121, :221-252).
_INT_FIELDS = ("Refine", "Degree", "Dimension", "Elements", "DoF", "NumEW")
_FLOAT_FIELDS = ("Area", "Volume", "BLength")
_STR_FIELDS = ("Creator", "File", "User")
# Timing fields: read under these labels (lapy/io.py:113-121) ...
_TIME_READ = {"Time(pre)": "TimePre", "Time(calcAB)": "TimeCalcAB",
              "Time(calcEW)": "TimeCalcEW"}
# ... and written under these (lapy/io.py:253-265). The written form has a
# capital "P" and " : " separators, so the reader does not find the timings
# again; both sides are kept as the reference has them.
_TIME_WRITE = (("TimePre", "Time(Pre)"), ("TimeCalcAB", "Time(calcAB)"),
               ("TimeCalcEW", "Time(calcEW)"))


def _open_text(filename, mode, what):
    try:
        return open(filename, mode)
    except OSError:
        logger.error("File %s %s", filename, what)
        raise


def read_vfunc(filename):
    """Read a PSOL vertex function: a ``Solution:`` line, then the values.

    Values may sit one per line, or on one line separated by ``,`` or ``;``;
    ``{}`` and ``()`` brackets are ignored (``lapy/io.py:10-53``).

    Raises ``OSError`` when the file cannot be read, ``ValueError`` when the
    ``Solution:`` line or the values are missing.
    """
    with _open_text(filename, "r", "not found or not readable") as f:
        lines = [ln.strip() for ln in f]
    try:
        lines.remove("Solution:")
    except ValueError:
        raise ValueError(f"Expected 'Solution:' marker in {filename}") from None
    body = [re.sub(r"[{()}]", "", ln) for ln in lines if ln]
    if not body:
        raise ValueError(f"No vertex function data found in {filename}")
    if len(body) == 1:
        body = re.split(r"[,;]", body[0])
    return [float(tok) for tok in body]


def write_vfunc(filename, vfunc):
    """Write a vertex function as PSOL text: ``Solution:`` then ``(v0,v1,...)``
    with no trailing newline (``lapy/io.py:287-310``)."""
    text = "Solution:\n(" + ",".join(np.asarray(vfunc).astype(str)) + ")"
    with _open_text(filename, "w", "not writable") as f:
        f.write(text)


def _braced_block(lines, i):
    """Collect the text between the first ``{`` at or after line ``i`` and the
    matching ``}``, with line contents concatenated (no separator added, as in
    ``lapy/io.py:128-136``). Returns the text and the index of the ``}`` line."""
    while "{" not in lines[i]:
        i += 1
    parts = []
    while True:
        parts.append(lines[i].strip())
        if "}" in lines[i]:
            break
        i += 1
    return "".join(parts).replace("{", "").replace("}", ""), i


def read_ev(filename):
    """Read an ``.ev`` file into a dict (``lapy/io.py:56-186``).

    Keys: the header fields present (``Refine``, ``Degree``, ``Dimension``,
    ``Elements``, ``DoF``, ``NumEW`` as int; ``Area``, ``Volume``, ``BLength``
    as float; ``EulerChar`` as int; ``Creator``, ``File``, ``User`` as str),
    ``Eigenvalues`` (float array), and, when an
    eigenvector block is present, ``EigenvectorsSize`` and ``Eigenvectors``
    (``(n, k)`` array, column ``j`` for eigenvalue ``j``). An eigenvector block
    whose length disagrees with ``sizes`` is reported on stdout and skipped.

    Two deliberate differences, both where the reference raises
    ``ValueError`` on a file its own writer produces: ``Creator``, ``File``
    and ``User`` are kept as strings (the reference casts them to int), and a
    one-line eigenvector block (``{ (...) }``, as written for a single
    eigenvector) is parsed (the reference leaves its parentheses in).

    Raises ``OSError`` when the file cannot be read.
    """
    with _open_text(filename, "r", "not readable") as f:
        lines = f.read().splitlines()
    d = {}
    i = 0
    while i < len(lines):
        line = lines[i].lstrip()
        key = line.split(":", 1)[0] if ":" in line else None
        if key in _INT_FIELDS or key in _FLOAT_FIELDS or key in _STR_FIELDS \
                or key == "EulerChar" or key in _TIME_READ:
            cast = float if key in _FLOAT_FIELDS else str if key in _STR_FIELDS else int
            d[_TIME_READ.get(key, key)] = cast(lines[i].split(":", 1)[1].strip())
            i += 1
        elif line.startswith("Eigenvalues"):
            text, i = _braced_block(lines, i + 1)
            d["Eigenvalues"] = np.array(text.split(";")).astype(float)
            i += 1
        elif line.startswith("Eigenvectors"):
            i += 1
            while not lines[i].strip().startswith("sizes"):
                i += 1
            sizes = np.array(lines[i].strip().split()[1:]).astype(int)
            d["EigenvectorsSize"] = sizes
            text, i = _braced_block(lines, i + 1)
            text = text.replace("(", " ").replace(")", " ")
            vals = np.array(re.split(r"[;,\s]+", text.strip())).astype(float)
            if len(vals) == sizes[0] * sizes[1]:
                # written column by column: k rows of n values
                d["Eigenvectors"] = vals.reshape(sizes[1], sizes[0]).T
            else:
                print(f"[Length of eigenvectors is not {sizes[0]} times {sizes[1]}.")
            i += 1
        else:
            i += 1
    return d


def write_ev(filename, d):
    """Write a dict of eigenvalues (and optionally eigenvectors and ShapeDNA
    header fields) as an ``.ev`` file (``lapy/io.py:189-284``).

    The text is the reference's byte for byte: header fields in their fixed
    order, each as `` Key: value``; eigenvalues as ``{ l0 ; l1 ; ... }``;
    eigenvectors as ``sizes: n k`` and one parenthesised, comma-separated
    column per line. Floats are written by ``str`` (shortest round-trip).

    Raises ``OSError`` when the file cannot be opened (before anything else)
    and ``ValueError`` when ``d`` has no ``Eigenvalues``.
    """
    f = _open_text(filename, "w", "not writable")
    with f:
        if "Eigenvalues" not in d:
            raise ValueError("ERROR: no Eigenvalues specified")
        out = []
        for key in _STR_FIELDS:
            if key in d:
                out.append(f" {key}: {d[key]}\n")
        for key in _INT_FIELDS:
            if key in d:
                out.append(f" {key}: {d[key]}\n")
        out.append("\n")
        for key in _FLOAT_FIELDS + ("EulerChar",):
            if key in d:
                out.append(f" {key}: {d[key]}\n")
        out.append("\n")
        for key, label in _TIME_WRITE:
            if key in d:
                out.append(f" {label} : {d[key]}\n")
        if all(key in d for key, _ in _TIME_WRITE):
            total = d["TimePre"] + d["TimeCalcAB"] + d["TimeCalcEW"]
            out.append(f" Time(total ) : {total}\n")
        out.append("\n")
        out.append("Eigenvalues:\n")
        out.append("{ " + " ; ".join(str(x) for x in d["Eigenvalues"]) + " }\n")
        out.append("\n")
        if "Eigenvectors" in d:
            evecs = d["Eigenvectors"]
            out.append("Eigenvectors:\n")
            out.append("sizes: " + " ".join(str(s) for s in evecs.shape) + "\n")
            out.append("\n")
            cols = ["(" + ",".join(str(x) for x in evecs[:, j]) + ")"
                    for j in range(evecs.shape[1])]
            out.append("{ " + " ;\n".join(cols) + " }\n")
        f.write("".join(out))
