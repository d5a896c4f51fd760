// HEALPix map files: FITS binary tables as written and read by HEALPix's
// write_Healpix_map_to_fits / read_Healpix_map_from_fits (used by the
// reference at source/cmb_gibbs.cpp:56-58, 92-104 and source/utils.cpp:26-52),
// and the NEST <-> RING index conversions.
//
// File layout (FITS standard, 2880-byte blocks of 80-character header cards,
// big-endian binary data): an empty primary HDU, then a BINTABLE extension
// with PIXTYPE = 'HEALPIX', ORDERING = 'RING' | 'NESTED', NSIDE, and the map
// in column 1 (further columns, e.g. Q and U, follow).  A column's TFORM is
// "<repeat><code>": the pixels run row by row, <repeat> per row.  Codes read:
// D (float64), E (float32), K (int64), J (int32), I (int16), B (uint8), with
// TSCALn / TZEROn applied.  Maps are written as float64 in rows of 1024.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "cmbl_internal.hpp"

namespace cmbl {
namespace {

constexpr size_t kBlock = 2880;
constexpr size_t kCard = 80;

struct Header {
    std::map<std::string, std::string> kv;  // keyword -> raw value text (quotes removed for strings)
    size_t data_offset = 0;                  // byte offset of the HDU's data
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    std::string str(const std::string& k) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? std::string() : it->second;
    }
    long num(const std::string& k, long dflt) const
    {
        auto it = kv.find(k);
        if (it == kv.end() || it->second.empty()) return dflt;
        return std::strtol(it->second.c_str(), nullptr, 10);
    }
    double real(const std::string& k, double dflt) const
    {
        auto it = kv.find(k);
        if (it == kv.end() || it->second.empty()) return dflt;
        std::string v = it->second;
        for (char& c : v)
            if (c == 'D' || c == 'd') c = 'E';  // FITS allows Fortran exponents
        return std::strtod(v.c_str(), nullptr);
    }
};

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Parse one header starting at byte `pos`; returns false on malformed input.
bool parse_header(const std::vector<char>& f, size_t pos, Header* h, std::string* err)
{
    for (size_t p = pos;; p += kCard) {
        if (p + kCard > f.size()) {
            *err = "FITS header without END";
            return false;
        }
        const std::string card(&f[p], kCard);
        const std::string key = trim(card.substr(0, 8));
        if (key == "END") {
            const size_t end = p + kCard;
            h->data_offset = (end + kBlock - 1) / kBlock * kBlock;
            return true;
        }
        if (card.size() < 10 || card[8] != '=' || key.empty()) continue;  // COMMENT, HISTORY, blank
        std::string val = card.substr(10);
        std::string v = trim(val);
        if (!v.empty() && v[0] == '\'') {
            // string value: up to the closing quote ('' is an escaped quote)
            std::string out;
            size_t i = 1;
            for (; i < v.size(); ++i) {
                if (v[i] == '\'') {
                    if (i + 1 < v.size() && v[i + 1] == '\'') {
                        out += '\'';
                        ++i;
                        continue;
                    }
                    break;
                }
                out += v[i];
            }
            h->kv[key] = trim(out);
        } else {
            const size_t slash = v.find('/');
            h->kv[key] = trim(slash == std::string::npos ? v : v.substr(0, slash));
        }
    }
}

size_t data_bytes(const Header& h)
{
    long bitpix = h.num("BITPIX", 8);
    long naxis = h.num("NAXIS", 0);
    if (naxis == 0) return 0;
    size_t n = 1;
    for (long i = 1; i <= naxis; ++i) n *= (size_t)h.num("NAXIS" + std::to_string(i), 0);
    n = n * (size_t)std::labs(bitpix) / 8 + (size_t)h.num("PCOUNT", 0);
    return n;
}

struct Column {
    long repeat = 1;
    char code = 'D';
    size_t offset = 0;  // byte offset inside a row
    double scale = 1.0, zero = 0.0;
};

size_t code_size(char c)
{
    switch (c) {
    case 'D': case 'K': return 8;
    case 'E': case 'J': return 4;
    case 'I': return 2;
    case 'B': case 'L': case 'A': case 'X': return 1;
    default: return 0;
    }
}

template <typename T>
T be(const char* p)
{
    unsigned char b[sizeof(T)];
    for (size_t i = 0; i < sizeof(T); ++i) b[i] = (unsigned char)p[sizeof(T) - 1 - i];
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

struct MapFile {
    std::vector<char> bytes;
    Header table;
    std::vector<Column> cols;
    long nside = 0, npix = 0, rows = 0, rowbytes = 0;
    bool nested = false;
};

int open_map(const char* path, MapFile* mf)
{
    if (!path) return fail(CMBL_ERR_ARG, "path is null");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(CMBL_ERR_ARG, std::string("cannot open ") + path);
    mf->bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    std::string err;
    Header h;
    size_t pos = 0;
    if (!parse_header(mf->bytes, 0, &h, &err)) return fail(CMBL_ERR_ARG, std::string(path) + ": " + err);
    if (h.str("SIMPLE") != "T") return fail(CMBL_ERR_ARG, std::string(path) + ": not a FITS file");
    // walk the extensions to the first binary table
    for (;;) {
        pos = h.data_offset + (data_bytes(h) + kBlock - 1) / kBlock * kBlock;
        if (pos >= mf->bytes.size()) return fail(CMBL_ERR_ARG, std::string(path) + ": no BINTABLE extension");
        h = Header();
        if (!parse_header(mf->bytes, pos, &h, &err)) return fail(CMBL_ERR_ARG, std::string(path) + ": " + err);
        if (h.str("XTENSION") == "BINTABLE") break;
    }
    mf->table = h;
    const std::string pix = h.str("PIXTYPE");
    if (!pix.empty() && pix != "HEALPIX") return fail(CMBL_ERR_ARG, std::string(path) + ": PIXTYPE is " + pix);
    if (h.str("INDXSCHM") == "EXPLICIT")
        return fail(CMBL_ERR_ARG, std::string(path) + ": explicitly indexed (partial-sky) maps are not supported");
    const std::string ord = h.str("ORDERING");
    if (ord == "NESTED" || ord == "NEST")
        mf->nested = true;
    else if (ord == "RING")
        mf->nested = false;
    else
        return fail(CMBL_ERR_ARG, std::string(path) + ": ORDERING must be RING or NESTED, got '" + ord + "'");
    mf->nside = h.num("NSIDE", 0);
    if (mf->nside < 1 || (mf->nside & (mf->nside - 1)) || mf->nside > (1L << 29))
        return fail(CMBL_ERR_ARG, std::string(path) + ": bad NSIDE");
    mf->npix = 12 * mf->nside * mf->nside;
    mf->rowbytes = h.num("NAXIS1", 0);
    mf->rows = h.num("NAXIS2", 0);
    const long nf = h.num("TFIELDS", 0);
    size_t off = 0;
    for (long c = 1; c <= nf; ++c) {
        const std::string t = h.str("TFORM" + std::to_string(c));
        size_t i = 0;
        while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i;
        Column col;
        col.repeat = i ? std::strtol(t.substr(0, i).c_str(), nullptr, 10) : 1;
        col.code = i < t.size() ? t[i] : '?';
        col.offset = off;
        col.scale = h.real("TSCAL" + std::to_string(c), 1.0);
        col.zero = h.real("TZERO" + std::to_string(c), 0.0);
        const size_t sz = code_size(col.code);
        if (!sz) return fail(CMBL_ERR_ARG, std::string(path) + ": unsupported TFORM " + t);
        off += sz * (size_t)col.repeat;
        mf->cols.push_back(col);
    }
    if ((long)off != mf->rowbytes) return fail(CMBL_ERR_ARG, std::string(path) + ": TFORMs do not add up to NAXIS1");
    if (mf->table.data_offset + (size_t)mf->rows * mf->rowbytes > mf->bytes.size())
        return fail(CMBL_ERR_ARG, std::string(path) + ": file shorter than its table");
    return CMBL_OK;
}

void put_card(std::string* h, const std::string& key, const std::string& value, const std::string& comment)
{
    std::string c = key;
    c.resize(8, ' ');
    c += "= ";
    std::string v = value;
    if (v.size() < 20 && (v.empty() || v[0] != '\'')) v = std::string(20 - v.size(), ' ') + v;
    c += v;
    if (!comment.empty()) c += " / " + comment;
    c.resize(kCard, ' ');
    *h += c;
}

std::string quoted(const std::string& s)
{
    std::string v = "'" + s;
    if (v.size() < 9) v.resize(9, ' ');
    return v + "'";
}

void end_header(std::string* h)
{
    std::string e = "END";
    e.resize(kCard, ' ');
    *h += e;
    h->resize((h->size() + kBlock - 1) / kBlock * kBlock, ' ');
}

}  // namespace
}  // namespace cmbl

using namespace cmbl;

extern "C" {

int cmbl_fits_map_info(const char* path, int* nside, int* nested, long* npix, int* ncolumns)
{
    MapFile mf;
    CMBL_TRY(open_map(path, &mf));
    if (nside) *nside = (int)mf.nside;
    if (nested) *nested = mf.nested ? 1 : 0;
    if (npix) *npix = mf.npix;
    if (ncolumns) *ncolumns = (int)mf.cols.size();
    return CMBL_OK;
}

int cmbl_fits_read_map(const char* path, int column, double* map, long npix)
{
    if (!map) return fail(CMBL_ERR_ARG, "map is null");
    MapFile mf;
    CMBL_TRY(open_map(path, &mf));
    if (column < 1 || column > (int)mf.cols.size())
        return fail(CMBL_ERR_ARG, "column " + std::to_string(column) + " out of range");
    if (npix != mf.npix) return fail(CMBL_ERR_ARG, "npix " + std::to_string(npix) + " != " + std::to_string(mf.npix));
    const Column& c = mf.cols[column - 1];
    if (c.repeat * mf.rows != mf.npix)
        return fail(CMBL_ERR_ARG, "column holds " + std::to_string(c.repeat * mf.rows) + " values for " +
                                      std::to_string(mf.npix) + " pixels");
    const size_t sz = code_size(c.code);
    long k = 0;
    for (long r = 0; r < mf.rows; ++r) {
        const char* row = &mf.bytes[mf.table.data_offset + (size_t)r * mf.rowbytes + c.offset];
        for (long i = 0; i < c.repeat; ++i, ++k) {
            const char* p = row + (size_t)i * sz;
            double v;
            switch (c.code) {
            case 'D': v = be<double>(p); break;
            case 'E': v = be<float>(p); break;
            case 'K': v = (double)be<int64_t>(p); break;
            case 'J': v = (double)be<int32_t>(p); break;
            case 'I': v = (double)be<int16_t>(p); break;
            default: v = (double)(unsigned char)*p; break;
            }
            map[k] = v * c.scale + c.zero;
        }
    }
    return CMBL_OK;
}

int cmbl_fits_write_map(const char* path, const double* map, int nside, int nested, int overwrite)
{
    if (!path || !map) return fail(CMBL_ERR_ARG, "null pointer");
    if (nside < 1 || (nside & (nside - 1))) return fail(CMBL_ERR_ARG, "nside must be a power of two");
    if (!overwrite) {
        std::ifstream probe(path);
        if (probe) return fail(CMBL_ERR_ARG, std::string(path) + " exists");
    }
    const long npix = 12L * nside * nside;
    const long repeat = npix < 1024 ? npix : 1024;  // npix = 12 * 4^k: 1024 divides it from nside 16 on
    const long rows = npix / repeat;
    std::string h;
    put_card(&h, "SIMPLE", "T", "file does conform to FITS standard");
    put_card(&h, "BITPIX", "8", "number of bits per data pixel");
    put_card(&h, "NAXIS", "0", "number of data axes");
    put_card(&h, "EXTEND", "T", "FITS dataset may contain extensions");
    end_header(&h);
    put_card(&h, "XTENSION", quoted("BINTABLE"), "binary table extension");
    put_card(&h, "BITPIX", "8", "8-bit bytes");
    put_card(&h, "NAXIS", "2", "2-dimensional binary table");
    put_card(&h, "NAXIS1", std::to_string(repeat * 8), "width of table in bytes");
    put_card(&h, "NAXIS2", std::to_string(rows), "number of rows in table");
    put_card(&h, "PCOUNT", "0", "size of special data area");
    put_card(&h, "GCOUNT", "1", "one data group");
    put_card(&h, "TFIELDS", "1", "number of fields in each row");
    put_card(&h, "TTYPE1", quoted("signal"), "label for field 1");
    put_card(&h, "TFORM1", quoted(std::to_string(repeat) + "D"), "data format of field: 8-byte DOUBLE");
    put_card(&h, "PIXTYPE", quoted("HEALPIX"), "HEALPIX pixelisation");
    put_card(&h, "ORDERING", quoted(nested ? "NESTED" : "RING"), "pixel ordering scheme");
    put_card(&h, "NSIDE", std::to_string(nside), "resolution parameter of HEALPIX");
    put_card(&h, "FIRSTPIX", "0", "first pixel # (0 based)");
    put_card(&h, "LASTPIX", std::to_string(npix - 1), "last pixel # (0 based)");
    put_card(&h, "INDXSCHM", quoted("IMPLICIT"), "indexing: IMPLICIT or EXPLICIT");
    end_header(&h);
    std::vector<char> data((size_t)npix * 8);
    for (long i = 0; i < npix; ++i) {
        unsigned char b[8];
        std::memcpy(b, &map[i], 8);
        for (int j = 0; j < 8; ++j) data[(size_t)i * 8 + j] = (char)b[7 - j];
    }
    data.resize((data.size() + kBlock - 1) / kBlock * kBlock, 0);
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return fail(CMBL_ERR_ARG, std::string("cannot write ") + path);
    out.write(h.data(), (std::streamsize)h.size());
    out.write(data.data(), (std::streamsize)data.size());
    if (!out) return fail(CMBL_ERR_ARG, std::string("write failed: ") + path);
    return CMBL_OK;
}

int cmbl_nest2ring(int nside, long ipnest, long* ipring)
{
    if (nside < 1 || (nside & (nside - 1)) || !ipring) return fail(CMBL_ERR_ARG, "bad nside or null output");
    if (ipnest < 0 || ipnest >= 12L * nside * nside) return fail(CMBL_ERR_ARG, "pixel out of range");
    *ipring = nest_to_ring(nside, ipnest);
    return CMBL_OK;
}

int cmbl_ring2nest(int nside, long ipring, long* ipnest)
{
    if (nside < 1 || (nside & (nside - 1)) || !ipnest) return fail(CMBL_ERR_ARG, "bad nside or null output");
    if (ipring < 0 || ipring >= 12L * nside * nside) return fail(CMBL_ERR_ARG, "pixel out of range");
    *ipnest = ring_to_nest(nside, ipring);
    return CMBL_OK;
}

}  // extern "C"
