// Reader of CASA AWProject convolution-function caches: directories of
// CFS_*.im (imaging) and WTCFS_*.im (weight) PagedImage<Complex> tables, the
// format AwConvolutionFunctionCache::open reads (awprojection.rs:920-1045).
//
// Every cell is a casacore table:
//   table.dat     AipsIO "Table" object: row count, data endianness, the table
//                 description with the image keywords ("coords", "miscinfo",
//                 ...), and the column set naming the data manager of the
//                 "map" column (table_control.rs:151-215, :352-720 of
//                 casa-tables restate casacore's layout);
//   table.f<N>    AipsIO header of that TiledCellStMan / TiledShapeStMan: cube
//                 and tile shapes, data file and offset (tiled_stman.rs:520-700);
//   table.f<N>_TSM<K>  the tiles, Fortran order within and across tiles, in
//                 the endianness the header records (tiled_stman.rs:1030-1080).
// AipsIO itself is canonical big-endian: an outer magic 0xbebebebe, then per
// object a length (counting from the length field), a type string and a
// version (casa-aipsio aipsio.rs:592-700).
//
// Host code only; the atlas upload and the per-sample selection kernel are in
// aw.hip.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <set>

#include "internal.hpp"

namespace chip {
namespace {

[[noreturn]] void cache_fail(const std::string &path, const std::string &msg) {
    fail(CHIP_ERR_CF_CACHE, path + ": " + msg);
}

// casacore DataType codes (casa-tables data_type.rs:10-40).
enum : int32_t {
    TpBool = 0, TpChar = 1, TpUChar = 2, TpShort = 3, TpUShort = 4, TpInt = 5, TpUInt = 6, TpFloat = 7,
    TpDouble = 8, TpComplex = 9, TpDComplex = 10, TpString = 11, TpTable = 12, TpArrayBool = 13,
    TpArrayChar = 14, TpArrayUChar = 15, TpArrayShort = 16, TpArrayUShort = 17, TpArrayInt = 18,
    TpArrayUInt = 19, TpArrayFloat = 20, TpArrayDouble = 21, TpArrayComplex = 22, TpArrayDComplex = 23,
    TpArrayString = 24, TpRecord = 25, TpInt64 = 29, TpArrayInt64 = 30
};

bool is_array_type(int32_t t) { return (t >= TpArrayBool && t <= TpArrayString) || t == TpArrayInt64; }

std::vector<uint8_t> read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) cache_fail(path, "cannot open file");
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n < 0) cache_fail(path, "cannot size file");
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> b(static_cast<size_t>(n));
    if (n > 0 && !f.read(reinterpret_cast<char *>(b.data()), n)) cache_fail(path, "cannot read file");
    return b;
}

// Canonical (big-endian) AipsIO reader over a file's bytes.
struct AipsIn {
    const std::vector<uint8_t> &b;
    std::string path;
    size_t pos = 0;
    std::vector<size_t> ends;  // end offset of every open object

    AipsIn(const std::vector<uint8_t> &bytes, std::string p) : b(bytes), path(std::move(p)) {}

    const uint8_t *take(size_t n) {
        const size_t lim = ends.empty() ? b.size() : ends.back();
        if (n > lim || pos > lim - n) cache_fail(path, "AipsIO object truncated");
        const uint8_t *r = b.data() + pos;
        pos += n;
        return r;
    }
    uint64_t be(size_t n) {
        const uint8_t *p = take(n);
        uint64_t v = 0;
        for (size_t k = 0; k < n; ++k) v = (v << 8) | p[k];
        return v;
    }
    uint8_t u8() { return static_cast<uint8_t>(be(1)); }
    bool boolean() { return u8() != 0; }
    int16_t i16() { return static_cast<int16_t>(be(2)); }
    uint16_t u16() { return static_cast<uint16_t>(be(2)); }
    uint32_t u32() { return static_cast<uint32_t>(be(4)); }
    int32_t i32() { return static_cast<int32_t>(u32()); }
    uint64_t u64() { return be(8); }
    int64_t i64() { return static_cast<int64_t>(u64()); }
    float f32() {
        const uint32_t v = u32();
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    }
    double f64() {
        const uint64_t v = u64();
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
    std::string str() {
        const uint32_t n = u32();
        const uint8_t *p = take(n);
        return std::string(reinterpret_cast<const char *>(p), n);
    }
    // Type of the next object without consuming it.
    std::string peek_type() {
        const size_t save = pos;
        if (ends.empty() && u32() != 0xbebebebeu) cache_fail(path, "no AipsIO magic value");
        (void)u32();
        std::string t = str();
        pos = save;
        return t;
    }
    // Open the next object, which must be of `type`; returns its version.
    uint32_t start(const std::string &type) {
        if (ends.empty() && u32() != 0xbebebebeu) cache_fail(path, "no AipsIO magic value");
        const size_t at = pos;
        const uint32_t len = u32();
        const size_t lim = ends.empty() ? b.size() : ends.back();
        if (len < 12 || at + len > lim) cache_fail(path, "invalid AipsIO object length");
        ends.push_back(at + len);
        const std::string t = str();
        if (t != type) cache_fail(path, "expected AipsIO object " + type + ", found " + t);
        return u32();
    }
    // Close the innermost object, skipping what was not read.
    void end() {
        pos = ends.back();
        ends.pop_back();
    }
};

// A keyword value: scalars, arrays (Fortran order) and sub-records.
struct Record;
struct Value {
    int32_t type = -1;
    bool b = false;
    int64_t i = 0;
    double d = 0.0, d_im = 0.0;
    std::string s;
    std::vector<double> dv;       // numeric arrays (real parts for complex)
    std::vector<std::string> sv;  // string arrays
    std::vector<uint64_t> shape;
    std::shared_ptr<Record> rec;
};

struct Record {
    std::vector<std::pair<std::string, Value>> fields;
    const Value *get(const std::string &name) const {
        for (const auto &f : fields)
            if (f.first == name) return &f.second;
        return nullptr;
    }
};

std::shared_ptr<Record> read_record_object(AipsIn &io);

// RecordDesc: field names and types; sub-record descriptions, array shapes and
// table-description names are skipped (table_control.rs:681-710).
std::vector<std::pair<std::string, int32_t>> read_record_desc(AipsIn &io) {
    io.start("RecordDesc");
    const int32_t n = io.i32();
    if (n < 0) cache_fail(io.path, "negative RecordDesc field count");
    std::vector<std::pair<std::string, int32_t>> d;
    for (int32_t k = 0; k < n; ++k) {
        std::string name = io.str();
        const int32_t t = io.i32();
        if (t == TpRecord) {
            (void)read_record_desc(io);
        } else if (is_array_type(t)) {
            io.start("IPosition");
            io.end();
        } else if (t == TpTable) {
            (void)io.str();
        }
        (void)io.str();  // comment
        d.emplace_back(std::move(name), t);
    }
    io.end();
    return d;
}

// Array<T> record field: "Array<...>" object, ndim, (origin before version 3),
// shape as ndim u32, element count, elements (table_control.rs:762-940).
Value read_array(AipsIn &io, int32_t t) {
    Value v;
    v.type = t;
    const uint32_t version = io.start(io.peek_type());
    const uint32_t ndim = io.u32();
    if (ndim > 16) cache_fail(io.path, "array of more than 16 dimensions");
    if (ndim == 0) {
        io.end();
        return v;
    }
    if (version < 3)
        for (uint32_t k = 0; k < ndim; ++k) (void)io.i32();
    uint64_t nel = 1;
    for (uint32_t k = 0; k < ndim; ++k) {
        v.shape.push_back(io.u32());
        nel *= v.shape.back();
    }
    const uint32_t count = io.u32();
    if (count > nel) cache_fail(io.path, "array element count exceeds its shape");
    if (t == TpArrayString) {
        for (uint32_t k = 0; k < count; ++k) v.sv.push_back(io.str());
    } else if (t == TpArrayBool) {
        // casacore stores Bool arrays one bit per element (AipsIO::put(uInt, const Bool*)).
        const uint8_t *p = io.take((count + 7) / 8);
        for (uint32_t k = 0; k < count; ++k) v.dv.push_back((p[k / 8] >> (k % 8)) & 1u);
    } else {
        for (uint32_t k = 0; k < count; ++k) {
            switch (t) {
                case TpArrayChar: v.dv.push_back(static_cast<int8_t>(io.u8())); break;
                case TpArrayUChar: v.dv.push_back(io.u8()); break;
                case TpArrayShort: v.dv.push_back(io.i16()); break;
                case TpArrayUShort: v.dv.push_back(io.u16()); break;
                case TpArrayInt: v.dv.push_back(io.i32()); break;
                case TpArrayUInt: v.dv.push_back(io.u32()); break;
                case TpArrayInt64: v.dv.push_back(static_cast<double>(io.i64())); break;
                case TpArrayFloat: v.dv.push_back(io.f32()); break;
                case TpArrayDouble: v.dv.push_back(io.f64()); break;
                case TpArrayComplex: v.dv.push_back(io.f32()); (void)io.f32(); break;
                case TpArrayDComplex: v.dv.push_back(io.f64()); (void)io.f64(); break;
                default: cache_fail(io.path, "unsupported array type " + std::to_string(t));
            }
        }
    }
    io.end();
    return v;
}

Value read_value(AipsIn &io, int32_t t) {
    Value v;
    v.type = t;
    switch (t) {
        case TpBool: v.b = io.boolean(); break;
        case TpChar: v.i = static_cast<int8_t>(io.u8()); break;
        case TpUChar: v.i = io.u8(); break;
        case TpShort: v.i = io.i16(); break;
        case TpUShort: v.i = io.u16(); break;
        case TpInt: v.i = io.i32(); break;
        case TpUInt: v.i = io.u32(); break;
        case TpInt64: v.i = io.i64(); break;
        case TpFloat: v.d = io.f32(); break;
        case TpDouble: v.d = io.f64(); break;
        case TpComplex: v.d = io.f32(); v.d_im = io.f32(); break;
        case TpDComplex: v.d = io.f64(); v.d_im = io.f64(); break;
        case TpString: v.s = io.str(); break;
        case TpTable: v.s = io.str(); break;
        case TpRecord: v.rec = read_record_object(io); break;
        default:
            if (is_array_type(t)) return read_array(io, t);
            cache_fail(io.path, "unsupported keyword type " + std::to_string(t));
    }
    return v;
}

// "TableRecord" (keyword sets) or "Record" (plain records): description,
// record type, then the values (table_control.rs:605-680).
std::shared_ptr<Record> read_record_object(AipsIn &io) {
    const std::string type = io.peek_type();
    if (type != "TableRecord" && type != "Record" && type != "RecordRep")
        cache_fail(io.path, "expected a record, found " + type);
    io.start(type);
    const auto desc = read_record_desc(io);
    (void)io.i32();  // record type
    auto r = std::make_shared<Record>();
    for (const auto &f : desc) r->fields.emplace_back(f.first, read_value(io, f.second));
    io.end();
    return r;
}

std::vector<int64_t> read_iposition(AipsIn &io) {
    io.start("IPosition");
    const uint32_t n = io.u32();
    if (n > 16) cache_fail(io.path, "IPosition of more than 16 axes");
    std::vector<int64_t> v(n);
    for (auto &x : v) x = io.i32();
    io.end();
    return v;
}

// What the CF reader needs of one PagedImage table.
struct ImageTable {
    std::shared_ptr<Record> keywords;
    std::vector<uint64_t> shape, tile;
    bool big_endian = true;
    std::string data_file;
    uint64_t data_offset = 0;
};

// table.dat: Table -> TableDesc (keywords, columns) -> ColumnSet (data
// managers, column -> data manager) (table_control.rs:151-215, :352-600).
void read_table_dat(const std::string &dir, ImageTable &img, std::string &dm_type, uint32_t &dm_seq) {
    const std::string path = dir + "/table.dat";
    const std::vector<uint8_t> bytes = read_file(path);
    AipsIn io(bytes, path);
    const uint32_t tv = io.start("Table");
    const uint64_t nrow = tv >= 3 ? io.u64() : io.u32();
    (void)io.u32();  // table-level endian format; tile data use the storage manager's flag
    const std::string kind = io.str();
    if (kind != "PlainTable") cache_fail(path, "expected a PlainTable, found " + kind);
    if (nrow != 1) cache_fail(path, "an image table has one row, found " + std::to_string(nrow));
    const uint32_t dv = io.start("TableDesc");
    (void)io.str(), (void)io.str(), (void)io.str();
    img.keywords = read_record_object(io);
    if (dv >= 2) (void)read_record_object(io);  // private keywords
    const uint32_t ncol = io.u32();
    if (ncol == 0 || ncol > 64) cache_fail(path, "unexpected column count " + std::to_string(ncol));
    std::vector<std::string> names(ncol);
    std::vector<bool> is_array(ncol);
    std::vector<int32_t> dtype(ncol);
    for (uint32_t c = 0; c < ncol; ++c) {
        (void)io.u32();
        const std::string cls = io.str();
        is_array[c] = cls.rfind("ArrayColumnDesc<", 0) == 0;
        const bool scalar = cls.rfind("ScalarColumnDesc<", 0) == 0;
        if (!is_array[c] && !scalar) cache_fail(path, "unsupported column class " + cls);
        (void)io.u32();
        names[c] = io.str();
        (void)io.str(), (void)io.str(), (void)io.str();  // comment, data manager type and group
        dtype[c] = io.i32();
        (void)io.i32(), (void)io.i32();  // option, ndim
        if (is_array[c]) (void)read_iposition(io);
        (void)io.u32();                       // maximum string length
        (void)read_record_object(io);         // column keywords
        (void)io.u32();
        if (is_array[c]) {
            (void)io.boolean();
        } else {
            switch (dtype[c]) {  // scalar default value
                case TpBool: case TpUChar: case TpChar: io.take(1); break;
                case TpShort: case TpUShort: io.take(2); break;
                case TpInt: case TpUInt: case TpFloat: io.take(4); break;
                case TpDouble: case TpInt64: case TpComplex: io.take(8); break;
                case TpDComplex: io.take(16); break;
                case TpString: (void)io.str(); break;
                default: cache_fail(path, "unsupported scalar column type");
            }
        }
    }
    io.end();
    const int32_t csv = io.i32();
    if (csv == -3) {
        (void)io.u64(), (void)io.i32(), (void)io.i32();
    } else {
        (void)io.u32();
    }
    (void)io.u32();  // sequence count
    const uint32_t ndm = io.u32();
    if (ndm == 0 || ndm > 64) cache_fail(path, "unexpected data manager count");
    std::vector<std::pair<std::string, uint32_t>> dms(ndm);
    for (auto &dm : dms) {
        dm.first = io.str();
        dm.second = io.u32();
    }
    int32_t map_col = -1;
    uint32_t map_seq = 0;
    for (uint32_t c = 0; c < ncol; ++c) {
        (void)io.u32();
        const std::string original = io.str();
        (void)io.u32();
        const uint32_t seq = io.u32();
        if (is_array[c] && io.boolean()) (void)read_iposition(io);
        if (names[c] == "map") map_col = static_cast<int32_t>(c), map_seq = seq;
    }
    io.end();
    if (map_col < 0) cache_fail(path, "no 'map' column: not a PagedImage");
    if (!is_array[map_col] || dtype[map_col] != TpComplex)
        cache_fail(path, "cannot open Complex32 image: the 'map' column is not a Complex array");
    for (const auto &dm : dms)
        if (dm.second == map_seq) dm_type = dm.first, dm_seq = dm.second;
    if (dm_type.empty()) cache_fail(path, "the 'map' column names no data manager");
}

// TiledStMan base object (tiled_stman.rs:620-740): endianness, data files
// and hypercubes.
struct TsmCube {
    std::vector<int64_t> shape, tile;
    int32_t file = -1;
    int64_t offset = 0;
};

std::vector<TsmCube> read_tsm_base(AipsIn &io, ImageTable &img, std::vector<int64_t> &file_seq) {
    const uint32_t v = io.start("TiledStMan");
    img.big_endian = v >= 2 ? io.boolean() : true;
    (void)io.u32();  // sequence number
    (void)(v >= 3 ? io.u64() : io.u32());
    const uint32_t ncol = io.u32();
    if (ncol > 64) cache_fail(io.path, "unexpected tiled column count");
    bool complex_col = false;
    for (uint32_t c = 0; c < ncol; ++c) complex_col = complex_col || io.i32() == TpComplex;
    if (ncol != 1 || !complex_col) cache_fail(io.path, "the tiled hypercolumn is not one Complex column");
    (void)io.str();
    (void)(v >= 3 ? io.u64() : io.u32());
    (void)io.u32();  // ndim
    const uint64_t nfiles = v >= 3 ? io.u64() : io.u32();
    if (nfiles > 4096) cache_fail(io.path, "unexpected data file count");
    file_seq.assign(nfiles, -1);
    for (uint64_t k = 0; k < nfiles; ++k) {
        if (!io.boolean()) continue;
        const uint32_t fv = io.u32();
        file_seq[k] = io.u32();
        (void)(fv >= 2 ? static_cast<uint64_t>(io.i64()) : io.u32());
    }
    const uint64_t ncubes = v >= 3 ? io.u64() : io.u32();
    if (ncubes > 1000000) cache_fail(io.path, "unexpected hypercube count");
    std::vector<TsmCube> cubes(ncubes);
    for (TsmCube &c : cubes) {
        const uint32_t cv = io.u32();
        (void)read_record_object(io);  // hypercube coordinate values
        (void)io.boolean();
        (void)io.u32();
        c.shape = read_iposition(io);
        c.tile = read_iposition(io);
        c.file = io.i32();
        c.offset = cv >= 2 ? io.i64() : static_cast<int64_t>(io.u32());
    }
    io.end();
    return cubes;
}

void use_cube(ImageTable &img, const std::vector<TsmCube> &cubes, const std::vector<int64_t> &file_seq,
              uint64_t index, const std::string &dir, uint32_t dm_seq, const std::string &path) {
    if (index >= cubes.size()) cache_fail(path, "row 0 has no hypercube");
    const TsmCube &c = cubes[index];
    if (c.shape.size() != c.tile.size() || c.shape.empty()) cache_fail(path, "cube and tile ranks differ");
    for (size_t d = 0; d < c.shape.size(); ++d) {
        if (c.shape[d] < 0 || c.tile[d] <= 0) cache_fail(path, "invalid cube or tile shape");
        img.shape.push_back(static_cast<uint64_t>(c.shape[d]));
        img.tile.push_back(static_cast<uint64_t>(c.tile[d]));
    }
    if (c.file < 0 || static_cast<uint64_t>(c.file) >= file_seq.size() || file_seq[c.file] < 0 || c.offset < 0)
        cache_fail(path, "the hypercube names no data file");
    img.data_file = dir + "/table.f" + std::to_string(dm_seq) + "_TSM" + std::to_string(file_seq[c.file]);
    img.data_offset = static_cast<uint64_t>(c.offset);
}

std::vector<uint32_t> read_block_u32(AipsIn &io) {
    io.start("Block");
    const uint32_t n = io.u32();
    if (n > (1u << 24)) cache_fail(io.path, "unexpected Block length");
    std::vector<uint32_t> v(n);
    for (auto &x : v) x = io.u32();
    io.end();
    return v;
}

ImageTable open_image(const std::string &dir) {
    ImageTable img;
    std::string dm_type;
    uint32_t dm_seq = 0;
    read_table_dat(dir, img, dm_type, dm_seq);
    const std::string hpath = dir + "/table.f" + std::to_string(dm_seq);
    const std::vector<uint8_t> bytes = read_file(hpath);
    AipsIn io(bytes, hpath);
    std::vector<int64_t> files;
    if (dm_type == "TiledCellStMan") {
        // default tile shape, then the base object; cube r holds row r
        // (tiled_stman.rs:563-570)
        io.start("TiledCellStMan");
        (void)read_iposition(io);
        const std::vector<TsmCube> cubes = read_tsm_base(io, img, files);
        io.end();
        use_cube(img, cubes, files, 0, dir, dm_seq, hpath);
    } else if (dm_type == "TiledShapeStMan") {
        // base object, default tile shape, used row-map length, then the row
        // map (last row of each interval), cube map and position map
        // (tiled_stman.rs:543-561): row 0 lies in the first interval
        io.start("TiledShapeStMan");
        const std::vector<TsmCube> cubes = read_tsm_base(io, img, files);
        (void)read_iposition(io);
        const uint32_t used = io.u32();
        const std::vector<uint32_t> rows = read_block_u32(io), cube_map = read_block_u32(io);
        io.end();
        if (used == 0 || rows.empty() || cube_map.empty()) cache_fail(hpath, "row 0 has no hypercube");
        use_cube(img, cubes, files, cube_map[0], dir, dm_seq, hpath);
    } else {
        cache_fail(dir, "unsupported image data manager " + dm_type);
    }
    return img;
}

// Pixels of the [nx, ny, 1, 1] cube as (x, y) with y fastest.
std::vector<float> read_pixels(const ImageTable &img, const std::string &dir) {
    const uint64_t nx = img.shape[0], ny = img.shape[1];
    const size_t nd = img.shape.size();
    std::vector<uint64_t> ntile(nd);
    uint64_t tiles = 1, tile_px = 1;
    for (size_t d = 0; d < nd; ++d) {
        ntile[d] = (img.shape[d] + img.tile[d] - 1) / img.tile[d];
        tiles *= ntile[d];
        tile_px *= img.tile[d];
    }
    const uint64_t tile_bytes = tile_px * 8;
    std::ifstream f(img.data_file, std::ios::binary);
    if (!f) cache_fail(dir, "cannot open tile data " + img.data_file);
    std::vector<float> out(2 * nx * ny);
    std::vector<uint8_t> t(tile_bytes);
    // Only tiles along axes 0 and 1 hold pixels of a [nx, ny, 1, 1] cube.
    for (uint64_t ty = 0; ty < ntile[1]; ++ty)
        for (uint64_t tx = 0; tx < ntile[0]; ++tx) {
            const uint64_t index = tx + ty * ntile[0];
            f.seekg(static_cast<std::streamoff>(img.data_offset + index * tile_bytes));
            if (!f.read(reinterpret_cast<char *>(t.data()), static_cast<std::streamsize>(tile_bytes)))
                cache_fail(dir, "tile data file is shorter than its hypercube");
            const uint64_t x0 = tx * img.tile[0], y0 = ty * img.tile[1];
            const uint64_t ex = std::min(img.tile[0], nx - x0), ey = std::min(img.tile[1], ny - y0);
            for (uint64_t j = 0; j < ey; ++j)
                for (uint64_t i = 0; i < ex; ++i)
                    for (int c = 0; c < 2; ++c) {
                        const uint8_t *p = t.data() + ((i + j * img.tile[0]) * 2 + c) * 4;
                        const uint32_t w = img.big_endian
                                               ? (uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3])
                                               : (uint32_t(p[3]) << 24 | uint32_t(p[2]) << 16 | uint32_t(p[1]) << 8 | p[0]);
                        float v;
                        std::memcpy(&v, &w, 4);
                        out[2 * ((x0 + i) * ny + (y0 + j)) + c] = v;
                    }
        }
    (void)tiles;
    return out;
}

const Record &sub_record(const Record &r, const std::string &name, const std::string &path) {
    const Value *v = r.get(name);
    if (!v || v->type != TpRecord || !v->rec) cache_fail(path, "missing " + name + " record");
    return *v->rec;
}

// Numeric vector of a coordinate field: an array, or a scalar as one value.
bool num_vec(const Record &r, const std::string &name, std::vector<double> &out) {
    const Value *v = r.get(name);
    if (!v) return false;
    if (v->type == TpDouble || v->type == TpFloat) {
        out = {v->d};
        return true;
    }
    if (v->type == TpInt || v->type == TpInt64 || v->type == TpShort || v->type == TpUInt) {
        out = {static_cast<double>(v->i)};
        return true;
    }
    if (is_array_type(v->type) && v->type != TpArrayString && v->type != TpArrayComplex &&
        v->type != TpArrayDComplex) {
        out = v->dv;
        return true;
    }
    return false;
}

std::vector<std::string> str_vec(const Record &r, const std::string &name) {
    const Value *v = r.get(name);
    if (!v) return {};
    if (v->type == TpString) return {v->s};
    if (v->type == TpArrayString) return v->sv;
    return {};
}

// casacore Stokes::StokesTypes codes by name.
int32_t stokes_code(const std::string &s) {
    static const char *names[] = {"Undefined", "I",  "Q",  "U",  "V",  "RR", "RL", "LR", "LL", "XX", "XY",
                                  "YX",        "YY", "RX", "RY", "LX", "LY", "XR", "XL", "YR", "YL", "PP",
                                  "PQ",        "QP", "QQ", "RCircular", "LCircular", "Linear", "Ptotal",
                                  "Plinear",   "PFtotal", "PFlinear", "Pangle"};
    for (int32_t k = 0; k < static_cast<int32_t>(sizeof(names) / sizeof(names[0])); ++k)
        if (s == names[k]) return k;
    return -1;
}

struct CoordOut {
    double uv[10];  // crval x2, crpix x2, cdelt x2, pc row-major 2x2
    int32_t polarization = 0;
    double frequency_hz = 0.0;
};

// The coordinate checks of validate_coordinates and spectral_reference_frequency
// (awprojection.rs:1448-1588): four pixel axes; a linear UU/VV coordinate in
// lambda with finite values, non-zero increments and an invertible PC; one
// Stokes value; a spectral coordinate with a finite positive reference value.
// Coordinates are the coords fields "coordinate<n>" (with coordinate_type) or
// "<type><n>" as casacore's CoordinateSystem::save writes them
// (coordinate_system.rs:212-280).
CoordOut parse_coordinates(const Record &coords, const std::string &path) {
    struct Entry { size_t index; std::string type; const Record *rec; };
    std::vector<Entry> modern, legacy;
    for (const auto &f : coords.fields) {
        if (f.second.type != TpRecord || !f.second.rec) continue;
        const std::string &n = f.first;
        auto index_of = [&](size_t from, size_t &out) {
            if (from >= n.size()) return false;
            for (size_t k = from; k < n.size(); ++k)
                if (n[k] < '0' || n[k] > '9') return false;
            out = std::stoul(n.substr(from));
            return true;
        };
        size_t idx = 0;
        if (n.rfind("coordinate", 0) == 0 && index_of(10, idx)) {
            const auto t = str_vec(*f.second.rec, "coordinate_type");
            modern.push_back({idx, t.empty() ? "" : t[0], f.second.rec.get()});
            continue;
        }
        static const char *prefixes[][2] = {{"direction", "Direction"}, {"spectral", "Spectral"},
                                            {"stokes", "Stokes"},       {"linear", "Linear"},
                                            {"tabular", "Tabular"}};
        for (const auto &p : prefixes)
            if (n.rfind(p[0], 0) == 0 && index_of(std::strlen(p[0]), idx)) {
                legacy.push_back({idx, p[1], f.second.rec.get()});
                break;
            }
    }
    std::vector<Entry> &es = modern.empty() ? legacy : modern;
    std::stable_sort(es.begin(), es.end(), [](const Entry &a, const Entry &b) { return a.index < b.index; });
    es.erase(std::unique(es.begin(), es.end(), [](const Entry &a, const Entry &b) { return a.index == b.index; }),
             es.end());
    size_t pixel_axes = 0;
    const Record *lin = nullptr, *stokes = nullptr, *spec = nullptr;
    for (const auto &e : es) {
        if (e.type == "Linear") {
            std::vector<double> cv;
            if (!num_vec(*e.rec, "crval", cv)) cache_fail(path, "linear coordinate has no crval");
            pixel_axes += cv.size();
            if (!lin) lin = e.rec;
        } else if (e.type == "Direction") {
            pixel_axes += 2;
        } else if (e.type == "Spectral" || e.type == "Stokes" || e.type == "Tabular") {
            pixel_axes += 1;
            if (e.type == "Spectral" && !spec) spec = e.rec;
            if (e.type == "Stokes" && !stokes) stokes = e.rec;
        } else {
            cache_fail(path, "unknown coordinate type " + e.type);
        }
    }
    if (pixel_axes != 4) cache_fail(path, "expected four CF coordinate axes, got " + std::to_string(pixel_axes));
    if (!lin) cache_fail(path, "missing UV linear coordinate");
    const auto axes = str_vec(*lin, "axes"), units = str_vec(*lin, "units");
    if (axes != std::vector<std::string>{"UU", "VV"} || units != std::vector<std::string>{"lambda", "lambda"})
        cache_fail(path, "expected UU/VV axes in lambda");
    std::vector<double> crval, crpix, cdelt, pc;
    if (!num_vec(*lin, "crval", crval) || !num_vec(*lin, "crpix", crpix) || !num_vec(*lin, "cdelt", cdelt))
        cache_fail(path, "UV linear coordinate lacks crval / crpix / cdelt");
    if (crval.size() != 2 || crpix.size() != 2 || cdelt.size() != 2)
        cache_fail(path, "expected a two-axis UU/VV coordinate");
    if (!num_vec(*lin, "pc", pc)) pc = {1.0, 0.0, 0.0, 1.0};
    if (pc.size() != 4) cache_fail(path, "linear pc matrix has " + std::to_string(pc.size()) + " elements");
    CoordOut o;
    // pc is a Matrix<Double> stored column-major: element (i, j) at i + 2 j.
    const double v[10] = {crval[0], crval[1], crpix[0], crpix[1], cdelt[0], cdelt[1], pc[0], pc[2], pc[1], pc[3]};
    for (int k = 0; k < 10; ++k) {
        if (!std::isfinite(v[k])) cache_fail(path, "UV affine coordinate contains a non-finite value");
        o.uv[k] = v[k];
    }
    if (o.uv[4] == 0.0 || o.uv[5] == 0.0) cache_fail(path, "UV coordinate increments must be non-zero");
    const double det = o.uv[6] * o.uv[9] - o.uv[7] * o.uv[8];
    if (!std::isfinite(det) || det == 0.0) cache_fail(path, "UV coordinate PC matrix must be invertible");
    if (!stokes) cache_fail(path, "missing Stokes coordinate");
    const Value *sv = stokes->get("stokes");
    std::vector<int32_t> codes;
    if (sv && (sv->type == TpArrayString || sv->type == TpString)) {
        for (const auto &s : str_vec(*stokes, "stokes")) {
            const int32_t c = stokes_code(s);
            if (c < 0) cache_fail(path, "invalid stokes name " + s);
            codes.push_back(c);
        }
    } else {
        std::vector<double> cv;
        if (!num_vec(*stokes, "stokes", cv)) cache_fail(path, "missing or invalid stokes");
        for (double c : cv) codes.push_back(static_cast<int32_t>(c));
    }
    if (codes.size() != 1) cache_fail(path, "expected one degenerate correlation coordinate");
    o.polarization = codes[0];
    if (!spec) cache_fail(path, "missing spectral coordinate");
    const Record *par = spec;
    if (const Value *w = spec->get("wcs"); w && w->type == TpRecord) par = w->rec.get();
    else if (const Value *t = spec->get("tabular"); t && t->type == TpRecord) par = t->rec.get();
    std::vector<double> fr;
    if (!num_vec(*par, "crval", fr) || fr.empty()) cache_fail(path, "spectral coordinate has no reference frequency");
    o.frequency_hz = fr[0];
    if (!(std::isfinite(o.frequency_hz) && o.frequency_hz > 0.0))
        cache_fail(path, "invalid spectral reference frequency");
    return o;
}

double misc_f64(const Record &m, const char *name, const std::string &path) {
    const Value *v = m.get(name);
    if (!v) cache_fail(path, std::string("missing miscinfo field ") + name);
    if (v->type != TpDouble && v->type != TpFloat)
        cache_fail(path, std::string("miscinfo field ") + name + " must be floating point");
    return v->d;
}

int32_t misc_i32(const Record &m, const char *name, const std::string &path) {
    const Value *v = m.get(name);
    if (!v) cache_fail(path, std::string("missing miscinfo field ") + name);
    if (v->type == TpInt) return static_cast<int32_t>(v->i);
    if (v->type == TpInt64) {
        if (v->i < INT32_MIN || v->i > INT32_MAX)
            cache_fail(path, std::string("miscinfo field ") + name + " is outside i32 range");
        return static_cast<int32_t>(v->i);
    }
    cache_fail(path, std::string("miscinfo field ") + name + " must be an integer");
}

bool misc_bool(const Record &m, const char *name, const std::string &path) {
    const Value *v = m.get(name);
    if (!v) cache_fail(path, std::string("missing miscinfo field ") + name);
    if (v->type != TpBool) cache_fail(path, std::string("miscinfo field ") + name + " must be Boolean");
    return v->b;
}

std::string misc_str(const Record &m, const char *name, const std::string &path) {
    const Value *v = m.get(name);
    if (!v) cache_fail(path, std::string("missing miscinfo field ") + name);
    if (v->type != TpString) cache_fail(path, std::string("miscinfo field ") + name + " must be a string");
    return v->s;
}

bool bits_eq(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }
uint64_t bits(double a) {
    uint64_t u;
    std::memcpy(&u, &a, 8);
    return u;
}

bool trimmed_empty(const std::string &s) {
    return s.find_first_not_of(" \t\n\r\f\v") == std::string::npos;
}

}  // namespace

// read_kernel_metadata and validate_kernel_metadata (awprojection.rs:1389-1446,
// :1590-1638).
CfKernelMeta cf_read_kernel_metadata(const std::string &dir, const std::string &relative, CfKey &key) {
    const ImageTable img = open_image(dir);
    const std::string &path = dir;
    if (img.shape.size() != 4 || img.shape[2] != 1 || img.shape[3] != 1 || img.shape[0] == 0 || img.shape[1] == 0)
        cache_fail(path, "expected non-empty [nx, ny, 1, 1] shape");
    const CoordOut co = parse_coordinates(sub_record(*img.keywords, "coords", path), path);
    const Record &misc = sub_record(*img.keywords, "miscinfo", path);
    key.frequency_hz = co.frequency_hz;
    key.w_value_lambda = misc_f64(misc, "WValue", path);
    key.mueller_element = misc_i32(misc, "MuellerElement", path);
    key.parallactic_angle_deg = misc_f64(misc, "ParallacticAngle", path);
    if (!std::isfinite(key.w_value_lambda)) cache_fail(path, "W value must be finite");
    if (!std::isfinite(key.parallactic_angle_deg)) cache_fail(path, "parallactic angle must be finite");
    CfKernelMeta m;
    m.path = dir;
    m.relative = relative;
    m.shape[0] = img.shape[0];
    m.shape[1] = img.shape[1];
    const double s = misc_f64(misc, "Sampling", path);
    if (!(std::isfinite(s) && s >= 1.0 && s == std::floor(s) && s <= 1.8e19))
        cache_fail(path, "miscinfo field Sampling must be a positive integral value");
    m.sampling = static_cast<uint64_t>(s);
    const int32_t xs = misc_i32(misc, "Xsupport", path), ys = misc_i32(misc, "Ysupport", path);
    if (xs < 0) cache_fail(path, "miscinfo field Xsupport must be non-negative");
    if (ys < 0) cache_fail(path, "miscinfo field Ysupport must be non-negative");
    if (xs == 0 || ys == 0) cache_fail(path, "Xsupport and Ysupport must be positive");
    m.x_support = static_cast<uint64_t>(xs);
    m.y_support = static_cast<uint64_t>(ys);
    std::memcpy(m.uv, co.uv, sizeof(m.uv));
    m.telescope = misc_str(misc, "TelescopeName", path);
    m.band = misc_str(misc, "BandName", path);
    m.diameter_m = misc_f64(misc, "Diameter", path);
    m.conj_freq_hz = misc_f64(misc, "ConjFreq", path);
    m.conj_poln = misc_i32(misc, "ConjPoln", path);
    m.polarization = co.polarization;
    m.w_increment = misc_f64(misc, "WIncr", path);
    m.rotationally_symmetric = misc_bool(misc, "OpCode", path);
    if (trimmed_empty(m.telescope) || trimmed_empty(m.band))
        cache_fail(path, "telescope and band names must be non-empty");
    if (!(std::isfinite(m.diameter_m) && m.diameter_m > 0.0)) cache_fail(path, "diameter must be finite and positive");
    if (!(std::isfinite(m.conj_freq_hz) && m.conj_freq_hz > 0.0))
        cache_fail(path, "conjugate frequency must be finite and positive");
    if (!(std::isfinite(m.w_increment) && m.w_increment >= 0.0))
        cache_fail(path, "W increment must be finite and non-negative");
    using u128 = unsigned __int128;
    if (u128(m.shape[0]) < u128(m.x_support) * m.sampling * 2 || u128(m.shape[1]) < u128(m.y_support) * m.sampling * 2)
        cache_fail(path, "shape cannot hold support " + std::to_string(m.x_support) + "x" +
                             std::to_string(m.y_support) + " at sampling " + std::to_string(m.sampling));
    return m;
}

// read_kernel_pixels (awprojection.rs:1980-2003): the first plane, (x, y) with
// y fastest, rejected when any value is not finite.
std::vector<float> cf_read_kernel_pixels(const CfKernelMeta &m) {
    const ImageTable img = open_image(m.path);
    if (img.shape.size() != 4 || img.shape[0] != m.shape[0] || img.shape[1] != m.shape[1] || img.shape[2] != 1 ||
        img.shape[3] != 1)
        cache_fail(m.path, "pixel shape changed after metadata indexing");
    std::vector<float> px = read_pixels(img, m.path);
    for (float v : px)
        if (!std::isfinite(v)) cache_fail(m.path, "pixel array contains non-finite complex values");
    return px;
}

namespace {

bool nearly_equal(double a, double b) {
    const double scale = std::max({std::fabs(a), std::fabs(b), 1.0});
    return std::fabs(a - b) <= scale * 32.0 * 2.220446049250313e-16;
}

// validate_pair (awprojection.rs:1640-1730).
void validate_pair(const CfKernelMeta &im, const CfKernelMeta &wt) {
    const std::pair<const char *, bool> checks[] = {
        {"sampling", im.sampling == wt.sampling},
        {"telescope", im.telescope == wt.telescope},
        {"band", im.band == wt.band},
        {"diameter", bits_eq(im.diameter_m, wt.diameter_m)},
        {"conjugate frequency", bits_eq(im.conj_freq_hz, wt.conj_freq_hz)},
        {"conjugate polarization", im.conj_poln == wt.conj_poln},
        {"polarization", im.polarization == wt.polarization},
        {"W increment", bits_eq(im.w_increment, wt.w_increment)},
        {"rotational symmetry", im.rotationally_symmetric == wt.rotationally_symmetric},
    };
    for (const auto &c : checks)
        if (!c.second)
            cache_fail(wt.path, std::string(c.first) + " does not match paired imaging cell " + im.path);
    bool same = bits_eq(im.uv[0], wt.uv[0]) && bits_eq(im.uv[1], wt.uv[1]);
    for (int k = 6; k < 10; ++k) same = same && bits_eq(im.uv[k], wt.uv[k]);
    for (int a = 0; a < 2 && same; ++a) {
        const double is = static_cast<double>(im.shape[a]), ws = static_cast<double>(wt.shape[a]);
        same = nearly_equal(im.uv[2 + a] / is, wt.uv[2 + a] / ws) && nearly_equal(im.uv[4 + a] * is, wt.uv[4 + a] * ws);
    }
    if (!same)
        cache_fail(wt.path, "weight CF coordinate does not cover the same UV world window as paired imaging cell " +
                                im.path);
}

bool key_less(const CfKey &a, const CfKey &b) {
    if (bits(a.frequency_hz) != bits(b.frequency_hz)) return bits(a.frequency_hz) < bits(b.frequency_hz);
    if (bits(a.w_value_lambda) != bits(b.w_value_lambda)) return bits(a.w_value_lambda) < bits(b.w_value_lambda);
    if (a.mueller_element != b.mueller_element) return a.mueller_element < b.mueller_element;
    return bits(a.parallactic_angle_deg) < bits(b.parallactic_angle_deg);
}

bool key_eq(const CfKey &a, const CfKey &b) { return !key_less(a, b) && !key_less(b, a); }

// Sorted unique values in total order (sorted_f64_bits, awprojection.rs:1974-1978).
std::vector<double> sorted_unique(const std::set<uint64_t> &b) {
    std::vector<double> v;
    for (uint64_t x : b) {
        double d;
        std::memcpy(&d, &x, 8);
        v.push_back(d);
    }
    std::sort(v.begin(), v.end(), [](double p, double q) {
        // f64::total_cmp for the non-NaN values a validated cache holds
        if (p != q) return p < q;
        return std::signbit(p) && !std::signbit(q);
    });
    return v;
}

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void bytes(const void *p, size_t n) {
        const uint8_t *c = static_cast<const uint8_t *>(p);
        for (size_t k = 0; k < n; ++k) {
            h ^= c[k];
            h *= 0x00000100000001b3ull;
        }
    }
    void u64(uint64_t v) {
        uint8_t le[8];
        for (int k = 0; k < 8; ++k) le[k] = static_cast<uint8_t>(v >> (8 * k));
        bytes(le, 8);
    }
    void i32(int32_t v) {
        uint8_t le[4];
        for (int k = 0; k < 4; ++k) le[k] = static_cast<uint8_t>(static_cast<uint32_t>(v) >> (8 * k));
        bytes(le, 4);
    }
    void text(const std::string &s) {
        u64(s.size());
        bytes(s.data(), s.size());
    }
    // StableFnv1a64::kernel_metadata (awprojection.rs:1945-1971)
    void meta(const CfKernelMeta &m) {
        text(m.relative);
        u64(m.shape[0]);
        u64(m.shape[1]);
        u64(m.sampling);
        u64(m.x_support);
        u64(m.y_support);
        for (int k = 0; k < 10; ++k) u64(bits(m.uv[k]));
        text(m.telescope);
        text(m.band);
        u64(bits(m.diameter_m));
        u64(bits(m.conj_freq_hz));
        i32(m.conj_poln);
        i32(m.polarization);
        u64(bits(m.w_increment));
        const uint8_t rs = m.rotationally_symmetric ? 1 : 0;
        bytes(&rs, 1);
    }
};

}  // namespace

// AwConvolutionFunctionCache::open (awprojection.rs:920-1045): index every
// CFS_*.im / WTCFS_*.im directory, pair them by key, validate pairs and axes,
// and derive the inventory and the metadata fingerprint.
void cf_cache_open(CfCache &c, const std::string &root) {
    DIR *d = opendir(root.c_str());
    if (!d) cache_fail(root, "cannot read cache directory");
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) names.emplace_back(e->d_name);
    closedir(d);
    std::sort(names.begin(), names.end());
    struct Partial { CfKey key; bool has_im = false, has_wt = false; CfKernelMeta im, wt; };
    std::vector<Partial> parts;
    size_t recognized = 0;
    auto ends_with = [](const std::string &s, const char *x) {
        const size_t n = std::strlen(x);
        return s.size() >= n && s.compare(s.size() - n, n, x) == 0;
    };
    for (const std::string &name : names) {
        const std::string path = root + "/" + name;
        struct stat st;
        if (stat(path.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) continue;
        bool weight;
        if (name.rfind("WTCFS_", 0) == 0 && ends_with(name, ".im")) weight = true;
        else if (name.rfind("CFS_", 0) == 0 && ends_with(name, ".im")) weight = false;
        else continue;
        ++recognized;
        CfKey key;
        CfKernelMeta meta = cf_read_kernel_metadata(path, name, key);
        auto it = std::find_if(parts.begin(), parts.end(), [&](const Partial &p) { return key_eq(p.key, key); });
        if (it == parts.end()) {
            parts.push_back(Partial{key});
            it = parts.end() - 1;
        }
        bool &has = weight ? it->has_wt : it->has_im;
        if (has)
            cache_fail(path, std::string("duplicate ") + (weight ? "weight" : "imaging") + " cell; first path is " +
                                 (weight ? it->wt.path : it->im.path));
        has = true;
        (weight ? it->wt : it->im) = std::move(meta);
    }
    if (recognized == 0) cache_fail(root, "no CFS_*.im or WTCFS_*.im entries found");
    std::sort(parts.begin(), parts.end(), [](const Partial &a, const Partial &b) { return key_less(a.key, b.key); });
    c = CfCache();
    c.root = root;
    for (const Partial &p : parts) {
        if (!p.has_im) cache_fail(root, "missing imaging CFS cell for a key");
        if (!p.has_wt) cache_fail(root, "missing weight WTCFS cell for a key");
        validate_pair(p.im, p.wt);
        c.keys.push_back(p.key);
        c.imaging.push_back(p.im);
        c.weight.push_back(p.wt);
    }
    // validate_cache_axes (awprojection.rs:1732-1790)
    const double wi = c.imaging[0].w_increment;
    for (const auto &m : c.imaging)
        if (!bits_eq(m.w_increment, wi)) cache_fail(root, "WIncr must be identical across one CF cache");
    std::set<uint64_t> fb, wb, pb;
    std::set<int32_t> mu;
    for (const CfKey &k : c.keys) {
        fb.insert(bits(k.frequency_hz));
        wb.insert(bits(k.w_value_lambda));
        pb.insert(bits(k.parallactic_angle_deg));
        mu.insert(k.mueller_element);
    }
    c.freqs = sorted_unique(fb);
    c.w_values = sorted_unique(wb);
    c.pas = sorted_unique(pb);
    c.muellers.assign(mu.begin(), mu.end());
    const uint64_t expected =
        static_cast<uint64_t>(c.freqs.size()) * c.w_values.size() * c.muellers.size() * c.pas.size();
    if (c.keys.size() != expected)
        cache_fail(root, "CF cache is not a complete frequency x W x Mueller x PA product: found " +
                             std::to_string(c.keys.size()) + " paired cells, expected " + std::to_string(expected));
    if (c.w_values.size() > 1 && wi == 0.0) cache_fail(root, "multiple W planes require a positive WIncr");
    if (wi > 0.0)
        for (size_t k = 0; k < c.w_values.size(); ++k) {
            const double want = static_cast<double>(k * k) / wi;
            if (std::fabs(c.w_values[k] - want) > std::max(std::fabs(want), 1.0) * 1.0e-10)
                cache_fail(root, "W plane " + std::to_string(k) + " has value " + std::to_string(c.w_values[k]) +
                                     ", expected CASA quadratic value from WIncr");
        }
    c.w_increment = wi;
    // identity_from_entries (awprojection.rs:1820-1900)
    Fnv h;
    h.text("casa-cf-cache-pagedimage-v1");
    for (size_t k = 0; k < c.keys.size(); ++k) {
        h.u64(bits(c.keys[k].frequency_hz));
        h.u64(bits(c.keys[k].w_value_lambda));
        h.i32(c.keys[k].mueller_element);
        h.u64(bits(c.keys[k].parallactic_angle_deg));
        h.meta(c.imaging[k]);
        h.meta(c.weight[k]);
    }
    c.fingerprint = h.h;
    // cell of every (frequency, W, Mueller, PA) index in inventory order
    const size_t nf = c.freqs.size(), nw = c.w_values.size(), nm = c.muellers.size(), np = c.pas.size();
    c.cell_of.assign(nf * nw * nm * np, 0);
    for (size_t k = 0; k < c.keys.size(); ++k) {
        const CfKey &y = c.keys[k];
        const size_t f = std::find_if(c.freqs.begin(), c.freqs.end(), [&](double v) { return bits_eq(v, y.frequency_hz); }) - c.freqs.begin();
        const size_t w = std::find_if(c.w_values.begin(), c.w_values.end(), [&](double v) { return bits_eq(v, y.w_value_lambda); }) - c.w_values.begin();
        const size_t m = std::find(c.muellers.begin(), c.muellers.end(), y.mueller_element) - c.muellers.begin();
        const size_t p = std::find_if(c.pas.begin(), c.pas.end(), [&](double v) { return bits_eq(v, y.parallactic_angle_deg); }) - c.pas.begin();
        c.cell_of[((f * nw + w) * nm + m) * np + p] = static_cast<uint32_t>(k);
    }
}

namespace {
double circular_degrees(double v) {
    double r = std::fmod(v + 180.0, 360.0);
    if (r < 0.0) r += 360.0;
    return r - 180.0;
}
}  // namespace

// nearest_key (awprojection.rs:1068-1104): Mueller exact, then the smallest
// relative frequency + relative W + PA/360 distance, ties by key order.
int64_t cf_nearest_key(const CfCache &c, double f, double w, int32_t mueller, double pa) {
    if (!(std::isfinite(f) && f > 0.0 && std::isfinite(w) && std::isfinite(pa))) return -1;
    int64_t best = -1;
    double bd = 0.0;
    const double fs = std::max(std::fabs(f), 1.0), ws = std::max(std::fabs(w), 1.0);
    for (size_t k = 0; k < c.keys.size(); ++k) {
        const CfKey &y = c.keys[k];
        if (y.mueller_element != mueller) continue;
        const double d = std::fabs((y.frequency_hz - f) / fs) + std::fabs((y.w_value_lambda - w) / ws) +
                         std::fabs(circular_degrees(y.parallactic_angle_deg - pa)) / 360.0;
        if (best < 0 || d < bd) best = static_cast<int64_t>(k), bd = d;  // keys are in key order
    }
    return best;
}

// select_key_for_sample / select_key_for_prediction_sample with CASA's
// CFBuffer axis semantics (awprojection.rs:1106-1300): the W plane
// round(sqrt(WIncr |w|)) clamped to the planes, the CF frequency nearest the
// requested one (the conjugate-beam map sqrt(2 f_ref^2 - f^2) when asked),
// the circularly nearest PA, ties to the smaller bit pattern; Mueller element
// m for w > 0 and 15 - m otherwise (prediction: the other way round).  Returns
// the cell index, or -1 where the reference returns None.
int64_t cf_select_key(const CfCache &c, double f, double f_ref, double w, int32_t mueller, double pa,
                      bool conjugate_beams, bool prediction) {
    if (!(std::isfinite(f) && f > 0.0 && std::isfinite(w) && std::isfinite(pa))) return -1;
    double req = f;
    if (!prediction) {
        if (!(std::isfinite(f_ref) && f_ref > 0.0)) return -1;
        if (conjugate_beams) {
            const double rad = 2.0 * f_ref * f_ref - f * f;
            if (!(std::isfinite(rad) && rad > 0.0)) return -1;
            req = std::sqrt(rad);
        }
    }
    const bool direct = prediction ? !(w > 0.0) : (w > 0.0);
    const int64_t m64 = direct ? static_cast<int64_t>(mueller) : 15 - static_cast<int64_t>(mueller);
    if (m64 < INT32_MIN || m64 > INT32_MAX) return -1;
    const auto mit = std::find(c.muellers.begin(), c.muellers.end(), static_cast<int32_t>(m64));
    if (mit == c.muellers.end()) return -1;
    const size_t nw = c.w_values.size();
    size_t wi = 0;
    if (nw > 1 && c.w_increment != 0.0) {
        double r = std::round(std::sqrt(c.w_increment * std::fabs(w)));
        r = std::min(std::max(r, 0.0), static_cast<double>(nw - 1));
        wi = static_cast<size_t>(r);
    }
    size_t fi = 0;
    for (size_t k = 1; k < c.freqs.size(); ++k) {
        const double a = std::fabs(c.freqs[k] - req), b = std::fabs(c.freqs[fi] - req);
        if (a < b || (a == b && bits(c.freqs[k]) < bits(c.freqs[fi]))) fi = k;
    }
    size_t pi = 0;
    for (size_t k = 1; k < c.pas.size(); ++k) {
        const double a = std::fabs(circular_degrees(c.pas[k] - pa)), b = std::fabs(circular_degrees(c.pas[pi] - pa));
        if (a < b || (a == b && bits(c.pas[k]) < bits(c.pas[pi]))) pi = k;
    }
    const size_t mi = static_cast<size_t>(mit - c.muellers.begin());
    return c.cell_of[((fi * nw + wi) * c.muellers.size() + mi) * c.pas.size() + pi];
}

}  // namespace chip
