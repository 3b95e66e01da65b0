// qkx_lime.cpp -- ILDG gauge files in LIME containers, read and written by
// our own record parser (no c-lime).
//
// Reference: qkxTM_MPI_getGaugeLime lib/read_conf_lime.cpp:111-264 and
// qkxTM_writeGaugeLime lib/write_conf_lime.cpp:49-199, which use the USQCD
// c-lime library (not part of the reference tree) for the records.
//
// LIME container (the published LIME format of c-lime): a sequence of
// records, each a 144-byte header followed by its payload, zero-padded to a
// multiple of 8 bytes.  Header, big-endian:
//   uint32 magic 0x456789ab | uint16 version 1 | uint16 flags (bit 15 MB
//   "message begin", bit 14 ME "message end") | uint64 payload bytes |
//   char type[128] (NUL-padded).
// ILDG payloads: "ildg-format" (XML with <precision> 32|64 and <lx>..<lt>)
// and "ildg-binary-data": big-endian reals in the order
// t, z, y, x, mu, row, col, re/im -- the reference's host layout
// U[lexic(t,z,y,x)][mu][row][col] once byte-swapped.
//
// Reading follows the reference: the records are scanned up to
// "ildg-binary-data", the precision comes from the last "ildg-format"
// record, and 64-bit data is accepted.  Beyond it: 32-bit data is accepted
// too (widened to double), the lattice in the XML and the payload size are
// checked against the requested lattice, and a truncated file is an error
// rather than a short read.  Writing produces the reference's three records
// ("ildg-format" with the same XML text, "xlf-info" with the caller's
// message, "ildg-binary-data" in 64-bit), so the files match byte for byte.
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "qkx_internal.h"
#include "qkx_core.h"

namespace {

constexpr uint32_t kLimeMagic = 0x456789abu;
constexpr size_t kHeader = 144;

uint64_t be_load(const unsigned char *p, int n)
{
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 8) | p[i];
    return v;
}
void be_store(unsigned char *p, uint64_t v, int n)
{
    for (int i = n - 1; i >= 0; i--) {
        p[i] = (unsigned char)(v & 0xff);
        v >>= 8;
    }
}

struct Record {
    bool mb, me;
    uint64_t bytes;
    std::string type;
    long long offset;   // file offset of the payload
};

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

// next record header at the current position; false at a clean end of file
int next_record(FILE *f, Record &r, bool &eof, const char *path)
{
    unsigned char h[kHeader];
    const size_t got = fread(h, 1, kHeader, f);
    if (got == 0 && feof(f)) {
        eof = true;
        return QKX_SUCCESS;
    }
    eof = false;
    if (got != kHeader) {
        qkx::set_error(std::string("LIME: truncated record header in ") + path);
        return QKX_ERR_ARG;
    }
    if (be_load(h, 4) != kLimeMagic) {
        qkx::set_error(std::string("LIME: bad magic number (not a LIME file?): ") + path);
        return QKX_ERR_ARG;
    }
    const uint64_t flags = be_load(h + 6, 2);
    r.mb = (flags >> 15) & 1;
    r.me = (flags >> 14) & 1;
    r.bytes = be_load(h + 8, 8);
    char type[129];
    memcpy(type, h + 16, 128);
    type[128] = '\0';
    r.type = type;
    r.offset = ftell(f);
    return QKX_SUCCESS;
}

int skip_payload(FILE *f, const Record &r)
{
    const uint64_t padded = (r.bytes + 7) / 8 * 8;
    if (fseek(f, (long)(r.offset + (long long)padded), SEEK_SET) != 0) {
        qkx::set_error("LIME: seek past a record failed");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

// value of <tag> ... </tag> in the XML text (trimmed); empty if absent
std::string xml_value(const std::string &xml, const char *tag)
{
    const std::string open = std::string("<") + tag + ">";
    const size_t a = xml.find(open);
    if (a == std::string::npos) return "";
    const size_t b = xml.find('<', a + open.size());
    std::string v = xml.substr(a + open.size(), b == std::string::npos ? std::string::npos : b - a - open.size());
    const size_t s = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
    return s == std::string::npos ? "" : v.substr(s, e - s + 1);
}

bool host_little_endian()
{
    const uint16_t one = 1;
    unsigned char c;
    memcpy(&c, &one, 1);
    return c == 1;
}

void swap8(unsigned char *p, size_t n)
{
    for (size_t i = 0; i < n; i++, p += 8)
        for (int k = 0; k < 4; k++) {
            const unsigned char t = p[k];
            p[k] = p[7 - k];
            p[7 - k] = t;
        }
}
void swap4(unsigned char *p, size_t n)
{
    for (size_t i = 0; i < n; i++, p += 4) {
        unsigned char t = p[0];
        p[0] = p[3];
        p[3] = t;
        t = p[1];
        p[1] = p[2];
        p[2] = t;
    }
}

int write_record(FILE *f, const char *type, const void *data, uint64_t bytes, bool mb, bool me)
{
    unsigned char h[kHeader];
    memset(h, 0, kHeader);
    be_store(h, kLimeMagic, 4);
    be_store(h + 4, 1, 2);
    be_store(h + 6, (mb ? 0x8000u : 0u) | (me ? 0x4000u : 0u), 2);
    be_store(h + 8, bytes, 8);
    strncpy((char *)h + 16, type, 128);
    if (fwrite(h, 1, kHeader, f) != kHeader) return -1;
    if (bytes && data && fwrite(data, 1, bytes, f) != bytes) return -1;
    static const unsigned char zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t pad = (8 - bytes % 8) % 8;
    if (pad && fwrite(zeros, 1, pad, f) != pad) return -1;
    return 0;
}

bool valid_lattice(const int L[4])
{
    for (int m = 0; m < 4; m++)
        if (L[m] <= 0) return false;
    return true;
}

} // namespace

extern "C" {

int qkx_lime_read_gauge(const char *path, const int L[4], double *U)
{
    QKX_CHECK_ARG(path && L && U && valid_lattice(L), "qkx_lime_read_gauge: bad argument");
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) {
        qkx::set_error(std::string("qkx_lime_read_gauge: cannot open ") + path);
        return QKX_ERR_ARG;
    }
    const uint64_t V = (uint64_t)L[0] * L[1] * L[2] * L[3];
    int precision = 0;
    std::string dims[4];
    Record r;
    bool eof = false;
    int rc;
    while (true) {
        if ((rc = next_record(fh.f, r, eof, path))) return rc;
        if (eof) {
            qkx::set_error(std::string("qkx_lime_read_gauge: no ildg-binary-data record in ") + path);
            return QKX_ERR_ARG;
        }
        if (r.type == "ildg-binary-data") break;
        if (r.type == "ildg-format") {
            std::string xml(r.bytes, '\0');
            if (r.bytes && fread(&xml[0], 1, r.bytes, fh.f) != r.bytes) {
                qkx::set_error("qkx_lime_read_gauge: truncated ildg-format record");
                return QKX_ERR_ARG;
            }
            precision = atoi(xml_value(xml, "precision").c_str());
            const char *tags[4] = {"lx", "ly", "lz", "lt"};
            for (int m = 0; m < 4; m++) dims[m] = xml_value(xml, tags[m]);
        }
        if ((rc = skip_payload(fh.f, r))) return rc;
    }
    if (precision != 64 && precision != 32) {
        qkx::set_error("qkx_lime_read_gauge: Unknown precision in configuration");
        return QKX_ERR_ARG;
    }
    for (int m = 0; m < 4; m++)
        if (!dims[m].empty() && atoi(dims[m].c_str()) != L[m]) {
            qkx::set_error("qkx_lime_read_gauge: the file's lattice differs from the requested one");
            return QKX_ERR_ARG;
        }
    const uint64_t w = precision / 8, n = V * 72;
    if (r.bytes != n * w) {
        qkx::set_error("qkx_lime_read_gauge: ildg-binary-data size does not match the lattice");
        return QKX_ERR_ARG;
    }
    const bool swap = host_little_endian();
    if (w == 8) {
        if (fread(U, 8, n, fh.f) != n) {
            qkx::set_error("qkx_lime_read_gauge: truncated ildg-binary-data");
            return QKX_ERR_ARG;
        }
        if (swap) swap8((unsigned char *)U, n);
    } else {
        // 32-bit: read in chunks and widen
        std::vector<float> buf(1 << 20);
        for (uint64_t done = 0; done < n;) {
            const uint64_t k = std::min<uint64_t>(buf.size(), n - done);
            if (fread(buf.data(), 4, k, fh.f) != k) {
                qkx::set_error("qkx_lime_read_gauge: truncated ildg-binary-data");
                return QKX_ERR_ARG;
            }
            if (swap) swap4((unsigned char *)buf.data(), k);
            for (uint64_t i = 0; i < k; i++) U[done + i] = (double)buf[i];
            done += k;
        }
    }
    return QKX_SUCCESS;
}

int qkx_lime_write_gauge(const char *path, const int L[4], const double *U, const char *message)
{
    QKX_CHECK_ARG(path && L && U && valid_lattice(L), "qkx_lime_write_gauge: bad argument");
    if (!message) message = "";
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) {
        qkx::set_error(std::string("qkx_lime_write_gauge: cannot open ") + path + " for writing");
        return QKX_ERR_ARG;
    }
    // the ildg-format text of lib/write_conf_lime.cpp:85-86
    char xml[2048];
    snprintf(xml, sizeof(xml),
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqkxTM.org/ildg\"\n"
             "            xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
             "            xsi:schemaLocation=\"http://www.lqkxTM.org/ildg filefmt.xsd\">\n  <version> 1.0 </version>\n"
             "  <field> su3gauge </field>\n  <precision> 64 </precision>\n  <lx> %d </lx>\n  <ly> %d </ly>\n"
             "  <lz> %d </lz>\n  <lt> %d </lt>\n</ildgFormat>",
             L[0], L[1], L[2], L[3]);
    const uint64_t V = (uint64_t)L[0] * L[1] * L[2] * L[3], n = V * 72;
    if (write_record(fh.f, "ildg-format", xml, strlen(xml), true, false) ||
        write_record(fh.f, "xlf-info", message, strlen(message), false, false) ||
        write_record(fh.f, "ildg-binary-data", nullptr, 0, false, true)) {
        qkx::set_error("qkx_lime_write_gauge: write failed");
        return QKX_ERR_ARG;
    }
    // the binary record was written with length 0 as a placeholder: patch
    // its length and append the payload (big-endian doubles, in chunks)
    const long hdr = ftell(fh.f) - (long)kHeader;
    unsigned char len[8];
    be_store(len, n * 8, 8);
    if (fseek(fh.f, hdr + 8, SEEK_SET) || fwrite(len, 1, 8, fh.f) != 8 || fseek(fh.f, 0, SEEK_END)) {
        qkx::set_error("qkx_lime_write_gauge: write failed");
        return QKX_ERR_ARG;
    }
    const bool swap = host_little_endian();
    std::vector<double> buf(1 << 20);
    for (uint64_t done = 0; done < n;) {
        const uint64_t k = std::min<uint64_t>(buf.size(), n - done);
        memcpy(buf.data(), U + done, k * 8);
        if (swap) swap8((unsigned char *)buf.data(), k);
        if (fwrite(buf.data(), 8, k, fh.f) != k) {
            qkx::set_error("qkx_lime_write_gauge: write failed");
            return QKX_ERR_ARG;
        }
        done += k;
    }
    if (fflush(fh.f)) {
        qkx::set_error("qkx_lime_write_gauge: write failed");
        return QKX_ERR_ARG;
    }
    return QKX_SUCCESS;
}

int qkx_gauge_read_lime(qkx_gauge *g, const char *path)
{
    QKX_CHECK_ARG(g && path, "qkx_gauge_read_lime: null argument");
    std::vector<double> U((size_t)g->V * 72);
    int rc = qkx_lime_read_gauge(path, g->L, U.data());
    if (rc) return rc;
    return qkx_gauge_upload(g, U.data());
}

int qkx_gauge_write_lime(const qkx_gauge *g, const char *path, const char *message)
{
    QKX_CHECK_ARG(g && path, "qkx_gauge_write_lime: null argument");
    std::vector<double> U((size_t)g->V * 72);
    int rc = qkx_gauge_download(g, U.data());
    if (rc) return rc;
    return qkx_lime_write_gauge(path, g->L, U.data(), message);
}

} // extern "C"
