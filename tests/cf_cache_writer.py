"""Writer of small CASA AWProject CF caches for the reader's tests.

A cache is a directory of CFS_*.im / WTCFS_*.im casacore PagedImage<Complex>
tables (awprojection.rs:920-1045 reads them).  Each table is written here the
way casacore lays one out, following the reference's own restatement of that
layout in casa-tables and casa-aipsio:
  table.dat            "Table" object: row count, endian flag, "PlainTable",
                       the "TableDesc" (keywords: coords, imageinfo, logtable,
                       miscinfo, units; one Complex array column "map") and the
                       column set (table_control.rs:963-1240);
  table.f0             the tiled storage manager's header (TiledCellStMan or
                       TiledShapeStMan around a "TiledStMan" base object of
                       version 1, 2 or 3; tiled_stman.rs:520-960);
  table.f0_TSM0        the tiles, Fortran order within and across tiles, big
                       or little endian as the header says;
  table.info           "Type = Image" / "SubType = PAGEDIMAGE".
AipsIO is canonical big-endian; the outermost object starts with the magic
0xbebebebe; each object is (length, type, version, payload) with the length
counted from the length field (aipsio.rs:592-640).

Test infrastructure only.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = 0xBEBEBEBE

# casacore DataType codes (casa-tables data_type.rs:10-40)
TP = dict(bool=0, uchar=2, short=3, int=5, uint=6, float=7, double=8, complex=9, dcomplex=10, string=11,
          table=12, array_bool=13, array_int=18, array_float=20, array_double=21, array_complex=22,
          array_string=24, record=25, int64=29)
ARRAY_TYPE_NAME = dict(array_int="Array<Int>", array_float="Array<float>", array_double="Array<double>",
                       array_string="Array<String>", array_complex="Array<void>", array_bool="Array<Bool>")


class AipsOut:
    """Canonical AipsIO writer."""

    def __init__(self):
        self.buf = bytearray()
        self.open = []

    def raw(self, fmt, *v):
        self.buf += struct.pack(">" + fmt, *v)

    def u8(self, v): self.raw("B", v)
    def boolean(self, v): self.raw("B", 1 if v else 0)
    def i16(self, v): self.raw("h", v)
    def i32(self, v): self.raw("i", v)
    def u32(self, v): self.raw("I", v)
    def i64(self, v): self.raw("q", v)
    def u64(self, v): self.raw("Q", v)
    def f32(self, v): self.raw("f", v)
    def f64(self, v): self.raw("d", v)

    def string(self, s):
        b = s.encode()
        self.u32(len(b))
        self.buf += b

    def start(self, kind, version):
        if not self.open:
            self.u32(MAGIC)
        self.open.append(len(self.buf))
        self.u32(MAGIC)  # length, patched by end()
        self.string(kind)
        self.u32(version)

    def end(self):
        at = self.open.pop()
        struct.pack_into(">I", self.buf, at, len(self.buf) - at)


# Typed keyword values: (type name, value[, shape]).  Records are lists of
# (name, typed value) pairs.
def D(x): return ("double", float(x))
def F(x): return ("float", float(x))
def I(x): return ("int", int(x))
def I64(x): return ("int64", int(x))
def B(x): return ("bool", bool(x))
def S(x): return ("string", str(x))
def T(x): return ("table", str(x))
def R(fields): return ("record", list(fields))
def AD(vals, shape=None): return ("array_double", [float(v) for v in np.ravel(vals, order="F")], shape or [np.size(vals)])
def AI(vals, shape=None): return ("array_int", [int(v) for v in np.ravel(vals, order="F")], shape or [np.size(vals)])
def AS(vals): return ("array_string", [str(v) for v in vals], [len(vals)])
def AB(vals): return ("array_bool", [bool(v) for v in vals], [len(vals)])


def write_record_desc(io, fields):
    io.start("RecordDesc", 2)
    io.i32(len(fields))
    for name, val in fields:
        io.string(name)
        io.i32(TP[val[0]])
        if val[0] == "record":
            write_record_desc(io, [])  # variable records describe no fields here
        elif val[0] == "table":
            io.string("")
        elif val[0].startswith("array"):
            write_iposition(io, val[2])
        io.string("")  # comment
    io.end()


def write_value(io, val, record_kind):
    t = val[0]
    if t == "double": io.f64(val[1])
    elif t == "float": io.f32(val[1])
    elif t == "int": io.i32(val[1])
    elif t == "int64": io.i64(val[1])
    elif t == "bool": io.boolean(val[1])
    elif t in ("string", "table"): io.string(val[1])
    elif t == "record": write_record(io, val[1], record_kind)
    elif t.startswith("array"):
        io.start(ARRAY_TYPE_NAME[t], 3)
        shape = val[2]
        io.u32(len(shape))
        for d in shape:
            io.u32(d)
        io.u32(len(val[1]))
        if t == "array_string":
            for s in val[1]:
                io.string(s)
        elif t == "array_bool":  # one bit per element, as casacore packs Bool arrays
            bits = bytearray((len(val[1]) + 7) // 8)
            for k, b in enumerate(val[1]):
                if b:
                    bits[k // 8] |= 1 << (k % 8)
            io.buf += bits
        else:
            f = {"array_double": io.f64, "array_float": io.f32, "array_int": io.i32}[t]
            for x in val[1]:
                f(x)
        io.end()
    else:
        raise ValueError(t)


def write_record(io, fields, kind="TableRecord"):
    io.start(kind, 1)
    write_record_desc(io, fields)
    io.i32(1)  # RecordInterface::Variable
    for _, val in fields:
        write_value(io, val, kind)
    io.end()


def write_iposition(io, vals):
    io.start("IPosition", 1)
    io.u32(len(vals))
    for v in vals:
        io.i32(int(v))
    io.end()


def coords_record(freq_hz, uv_crval, uv_crpix, uv_cdelt, stokes="RR", pc=((1.0, 0.0), (0.0, 1.0)),
                  stokes_as_codes=False):
    """CoordinateSystem::save layout: linear0 (UU/VV), stokes1, spectral2,
    axis maps and an obsinfo block (coordinate_system.rs:156-210)."""
    codes = {"I": 1, "Q": 2, "U": 3, "V": 4, "RR": 5, "RL": 6, "LR": 7, "LL": 8, "XX": 9, "XY": 10, "YX": 11,
             "YY": 12}
    pcm = np.array(pc, dtype=np.float64)
    return R([
        ("obsinfo", R([("telescope", S("EVLA")), ("observer", S("")),
                       ("obsdate", R([("type", S("epoch")), ("refer", S("UTC")),
                                      ("m0", R([("value", D(0.0)), ("unit", S("d"))]))])),
                       ("pointingcenter", R([("value", AD([0.0, 0.0])), ("initial", B(True))]))])),
        ("linear0", R([("coordinate_type", S("Linear")), ("naxes", I(2)), ("crval", AD(uv_crval)),
                       ("cdelt", AD(uv_cdelt)), ("crpix", AD(uv_crpix)), ("pc", AD(pcm, [2, 2])),
                       ("axes", AS(["UU", "VV"])), ("units", AS(["lambda", "lambda"]))])),
        ("worldmap0", AI([0, 1])), ("pixelmap0", AI([0, 1])), ("pixelreplace0", AD([0.0, 0.0])),
        ("stokes1", R([("axes", AS(["Stokes"])),
                       ("stokes", AI([codes[stokes]]) if stokes_as_codes else AS([stokes])),
                       ("crval", AD([codes[stokes]])), ("crpix", AD([0.0])), ("cdelt", AD([1.0])),
                       ("pc", AD([[1.0]], [1, 1]))])),
        ("worldmap1", AI([2])), ("pixelmap1", AI([2])),
        ("spectral2", R([("version", I(2)), ("system", S("LSRK")), ("restfreq", D(freq_hz)),
                         ("restfreqs", AD([freq_hz])), ("velType", I(0)), ("nativeType", I(0)),
                         ("velUnit", S("km/s")), ("waveUnit", S("mm")), ("formatUnit", S("")),
                         ("unit", S("Hz")), ("name", S("Frequency")),
                         ("wcs", R([("crval", D(freq_hz)), ("crpix", D(0.0)), ("cdelt", D(1.0)),
                                    ("pc", D(1.0)), ("ctype", S("FREQ"))]))])),
        ("worldmap2", AI([3])), ("pixelmap2", AI([3])),
    ])


def misc_record(**kw):
    """miscinfo of one CF cell as CASA's AWConvFunc writes it (awprojection.rs:3095-3110)."""
    f = dict(BandName=S("EVLA_L"), ConjFreq=D(kw.get("conj_freq", 1.0e9)), ConjPoln=I(8), Diameter=D(25.0),
             MuellerElement=I(kw.get("mueller", 0)), Name=S(kw.get("name", "cf")), OpCode=B(False),
             ParallacticAngle=D(kw.get("pa", 30.0)), Sampling=D(kw.get("sampling", 2.0)),
             TelescopeName=S("EVLA"), WIncr=D(kw.get("w_incr", 0.5)), WValue=D(kw.get("w", 0.0)),
             Xsupport=I(kw.get("x_support", 2)), Ysupport=I(kw.get("y_support", 2)))
    for k, v in kw.get("override", {}).items():
        if v is None:
            f.pop(k, None)
        else:
            f[k] = v
    return R(sorted(f.items()))


def write_paged_image(path, pixels, coords, misc, tile=None, big_endian=False, tsm_version=None,
                      manager="TiledCellStMan"):
    """One PagedImage<Complex> of shape [nx, ny, 1, 1] holding `pixels`
    (complex64, indexed [x, y])."""
    pixels = np.asarray(pixels, np.complex64)
    nx, ny = pixels.shape
    shape = [nx, ny, 1, 1]
    tile = list(tile or [nx, ny, 1, 1])
    os.makedirs(path, exist_ok=True)
    keywords = [("coords", coords), ("imageinfo", R([("imagetype", S("Intensity")), ("objectname", S(""))])),
                ("logtable", T("logtable")), ("miscinfo", misc), ("units", S(""))]

    io = AipsOut()
    io.start("Table", 2)
    io.u32(1)
    io.u32(0 if big_endian else 1)
    io.string("PlainTable")
    io.start("TableDesc", 2)
    io.string(""); io.string(""); io.string("")
    write_record(io, keywords)
    write_record(io, [])
    io.u32(1)
    io.u32(1)
    io.string("ArrayColumnDesc<Complex   >")
    io.u32(1)
    io.string("map"); io.string(""); io.string(manager); io.string(manager)
    io.i32(TP["complex"]); io.i32(0); io.i32(4)
    write_iposition(io, [])
    io.u32(0)
    write_record(io, [])
    io.u32(1)
    io.boolean(False)
    io.end()
    io.i32(-2); io.u32(1); io.u32(1)
    io.u32(1); io.string(manager); io.u32(0)
    io.u32(2); io.string("map"); io.u32(1); io.u32(0); io.boolean(False)
    io.u32(0)
    io.end()
    with open(os.path.join(path, "table.dat"), "wb") as fh:
        fh.write(io.buf)
    with open(os.path.join(path, "table.info"), "w") as fh:
        fh.write("Type = Image\nSubType = PAGEDIMAGE\n")

    # tiles, Fortran order within and across tiles
    ntile = [-(-shape[d] // tile[d]) for d in range(4)]
    tile_px = int(np.prod(tile))
    cube = np.zeros((nx, ny), np.complex64)
    cube[:, :] = pixels
    data = bytearray()
    dt = np.dtype(">f4" if big_endian else "<f4")
    for ty in range(ntile[1]):
        for tx in range(ntile[0]):
            t = np.zeros((tile[0], tile[1]), np.complex64)
            x0, y0 = tx * tile[0], ty * tile[1]
            blk = cube[x0:x0 + tile[0], y0:y0 + tile[1]]
            t[:blk.shape[0], :blk.shape[1]] = blk
            full = np.zeros(tile_px, np.complex64)
            full[:tile[0] * tile[1]] = t.ravel(order="F")
            pairs = np.empty(2 * tile_px, np.float32)
            pairs[0::2] = full.real
            pairs[1::2] = full.imag
            data += pairs.astype(dt).tobytes()
    with open(os.path.join(path, "table.f0_TSM0"), "wb") as fh:
        fh.write(data)

    v = tsm_version or (1 if big_endian else 2)
    h = AipsOut()
    h.start(manager, 1)
    if manager == "TiledCellStMan":
        write_iposition(h, tile)
    h.start("TiledStMan", v)
    if v >= 2:
        h.boolean(big_endian)
    h.u32(0)
    (h.u64 if v >= 3 else h.u32)(1)
    h.u32(1); h.i32(TP["complex"])
    h.string("map")
    (h.u64 if v >= 3 else h.u32)(0)
    h.u32(4)
    (h.u64 if v >= 3 else h.u32)(1)
    h.boolean(True); h.u32(1); h.u32(0); h.u32(len(data))
    cubes = [] if manager == "TiledCellStMan" else [([0, 0, 0, 0], tile, 0)]  # TiledShapeStMan's empty cube 0
    cubes.append((shape, tile, 0))
    (h.u64 if v >= 3 else h.u32)(len(cubes))
    for cshape, ctile, off in cubes:
        h.u32(1)
        write_record(h, [], "Record")
        h.boolean(False)
        h.u32(4)
        write_iposition(h, cshape)
        write_iposition(h, ctile)
        h.i32(0)
        h.u32(off)
    h.end()
    if manager == "TiledShapeStMan":
        write_iposition(h, tile)
        h.u32(1)
        for block in ([0], [len(cubes) - 1], [0]):  # row map, cube map, position map
            h.start("Block", 1)
            h.u32(len(block))
            for b in block:
                h.u32(b)
            h.end()
    h.end()
    with open(os.path.join(path, "table.f0"), "wb") as fh:
        fh.write(h.buf)
    return path


def write_test_cell(root, name, frequency_hz, w_value_lambda, weight, uv_reference_value=(0.0, 0.0),
                    mueller_element=0, conjugate_frequency_hz=None, pa=30.0, w_incr=0.5, pixels=None,
                    sampling=2.0, support=None, crpix=None, cdelt=None, **layout):
    """The reference test cell (awprojection.rs:3045-3114): imaging 16x16 of
    3-1j, reference pixel 8, increment (-2, 2), support 2; weight 32x32 of
    7+2j, reference pixel 16, increment (-1, 1), support 4; sampling 2,
    WIncr 0.5, PA 30, ConjFreq 2 GHz - f."""
    uv_crpix, uv_cdelt = crpix, cdelt
    if weight:
        crpix, cdelt, n, val, sup = (16.0, 16.0), (-1.0, 1.0), 32, 7.0 + 2.0j, 4
    else:
        crpix, cdelt, n, val, sup = (8.0, 8.0), (-2.0, 2.0), 16, 3.0 - 1.0j, 2
    if pixels is None:
        pixels = np.full((n, n), val, np.complex64)
    if support is not None:
        sup = support
    if uv_crpix is not None:
        crpix, cdelt = tuple(uv_crpix), tuple(uv_cdelt)
    conj = 2.0e9 - frequency_hz if conjugate_frequency_hz is None else conjugate_frequency_hz
    misc = misc_record(conj_freq=conj, mueller=mueller_element, name=name, pa=pa, w=w_value_lambda, w_incr=w_incr,
                       x_support=sup, y_support=sup, sampling=sampling, override=layout.pop("override", {}))
    coords = coords_record(frequency_hz, uv_reference_value, crpix, cdelt,
                           stokes_as_codes=layout.pop("stokes_as_codes", False))
    return write_paged_image(os.path.join(root, name), pixels, coords, misc, **layout)
