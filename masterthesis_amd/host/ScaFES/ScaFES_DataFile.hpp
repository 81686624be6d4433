// ScaFES_DataFile.hpp -- snapshot / initial-state files of the host API.
//
// Same file layout as ScaFES::DataFile (src/ScaFES_DataFileApi.hpp:392-585):
// a NetCDF dataset <name>.nc with dimensions nNodes_0 .. nNodes_{DIM-1} and
// the unlimited "time" (defined in that order), a scalar NC_INT variable
// "version" = 3, and one NC_DOUBLE variable per written data field with
// dimensions (time, nNodes_{DIM-1}, .., nNodes_0), i.e. x fastest.  Each
// write() appends one record; variables not written in that record hold the
// NetCDF fill value.  read() takes record 0 of each named variable, as the
// reference's nc_get_var_double into a one-record buffer does.
//
// The reference's serial build passes NC_FORMAT_NETCDF4 (= 3) as the mode
// of nc_create, which selects no format bit, so its files are NetCDF
// classic.  This engine writes the classic format itself (CDF-1; CDF-2
// when offsets need 64 bits) from the published format specification, so
// no libnetcdf is needed; the reader accepts CDF-1 and CDF-2 files with
// attributes and byte/short/int/float/double variables (converted to
// double, as nc_get_var_double converts).
//
// WriteHowOften values are the reference's (src/ScaFES_DataFile.hpp:24-32).
#ifndef SCAFES_HOST_DATAFILE_HPP_
#define SCAFES_HOST_DATAFILE_HPP_

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ScaFES_Ntuple.hpp"

namespace ScaFES {

enum class WriteHowOften {
    NEVER = 1,
    AT_START = 2,
    AT_END = 3,
    AT_START_AND_END = 4,
    LIKE_GIVEN_AT_CL = 8,
    ALWAYS = 16,
};

namespace netcdf {

const uint32_t kDimension = 0x0A, kVariable = 0x0B, kAttribute = 0x0C;
const uint32_t kByte = 1, kChar = 2, kShort = 3, kInt = 4, kFloat = 5, kDouble = 6;
const double kFillDouble = 9.9692099683868690e+36;  // NC_FILL_DOUBLE
const int kDataFileVersion = 3;                      // DATA_FILE_VERSION

inline void put32(std::string& b, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) b.push_back((char)((v >> s) & 0xFF));
}
inline void put64(std::string& b, uint64_t v)
{
    for (int s = 56; s >= 0; s -= 8) b.push_back((char)((v >> s) & 0xFF));
}
inline void putName(std::string& b, const std::string& name)
{
    put32(b, (uint32_t)name.size());
    b += name;
    while (b.size() % 4) b.push_back('\0');
}
inline uint64_t beBits(const unsigned char* p, int n)
{
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
    return v;
}
inline void putDoublesBE(std::string& b, const double* v, std::size_t n)
{
    for (std::size_t i = 0; i < n; ++i) {
        uint64_t u;
        std::memcpy(&u, &v[i], 8);
        put64(b, u);
    }
}
inline std::size_t typeSize(uint32_t t)
{
    switch (t) {
    case kByte: case kChar: return 1;
    case kShort: return 2;
    case kInt: case kFloat: return 4;
    case kDouble: return 8;
    default: throw std::runtime_error("NetCDF: unknown data type");
    }
}

// Header of a classic-format file, as far as reading variables needs it.
struct Header {
    struct Dim {
        std::string name;
        uint64_t len;
    };
    struct Var {
        std::string name;
        std::vector<uint32_t> dims;
        uint32_t type;
        uint64_t vsize, begin;
    };
    int version = 1;
    uint64_t numrecs = 0;
    std::vector<Dim> dims;
    std::vector<Var> vars;
    bool isRecordVar(const Var& v) const { return !v.dims.empty() && dims.at(v.dims[0]).len == 0; }
    uint64_t recsize() const
    {
        uint64_t s = 0;
        int nrec = 0;
        for (const Var& v : vars)
            if (isRecordVar(v)) {
                s += v.vsize;
                ++nrec;
            }
        // a lone record variable is not padded (classic format rule)
        if (nrec == 1)
            for (const Var& v : vars)
                if (isRecordVar(v)) {
                    uint64_t n = typeSize(v.type);
                    for (std::size_t d = 1; d < v.dims.size(); ++d) n *= dims.at(v.dims[d]).len;
                    return n;
                }
        return s;
    }
};

class Reader {
public:
    explicit Reader(const std::string& path) : mPath(path)
    {
        std::ifstream in(path.c_str(), std::ios::binary);
        if (!in) throw std::runtime_error("NetCDF: cannot open " + path);
        mBytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        parse();
    }
    const Header& header() const { return mH; }
    // Record `rec` (ignored for fixed-size variables) of variable `name`,
    // converted to double.
    std::vector<double> read(const std::string& name, uint64_t rec = 0) const
    {
        const Header::Var* v = 0;
        for (const Header::Var& x : mH.vars)
            if (x.name == name) v = &x;
        if (!v) throw std::runtime_error("NetCDF: variable " + name + " not found in " + mPath);
        uint64_t count = 1;
        const bool isRec = mH.isRecordVar(*v);
        for (std::size_t d = isRec ? 1 : 0; d < v->dims.size(); ++d) count *= mH.dims.at(v->dims[d]).len;
        if (isRec && rec >= mH.numrecs) throw std::runtime_error("NetCDF: record out of range in " + mPath);
        const uint64_t off = v->begin + (isRec ? rec * mH.recsize() : 0);
        const std::size_t ts = typeSize(v->type);
        if (off + count * ts > mBytes.size()) throw std::runtime_error("NetCDF: file truncated: " + mPath);
        std::vector<double> out((std::size_t)count);
        const unsigned char* p = (const unsigned char*)mBytes.data() + off;
        for (uint64_t i = 0; i < count; ++i, p += ts) {
            const uint64_t u = beBits(p, (int)ts);
            switch (v->type) {
            case kByte: out[i] = (double)(int8_t)u; break;
            case kChar: out[i] = (double)(uint8_t)u; break;
            case kShort: out[i] = (double)(int16_t)u; break;
            case kInt: out[i] = (double)(int32_t)u; break;
            case kFloat: {
                const uint32_t w = (uint32_t)u;
                float f;
                std::memcpy(&f, &w, 4);
                out[i] = (double)f;
                break;
            }
            default: {
                double d;
                std::memcpy(&d, &u, 8);
                out[i] = d;
            }
            }
        }
        return out;
    }

private:
    uint64_t take(int n)
    {
        if (mPos + n > mBytes.size()) throw std::runtime_error("NetCDF: header truncated: " + mPath);
        const uint64_t v = beBits((const unsigned char*)mBytes.data() + mPos, n);
        mPos += n;
        return v;
    }
    std::string name()
    {
        const uint64_t n = take(4);
        if (mPos + n > mBytes.size()) throw std::runtime_error("NetCDF: header truncated: " + mPath);
        std::string s(mBytes.data() + mPos, (std::size_t)n);
        mPos += (std::size_t)((n + 3) / 4 * 4);
        return s;
    }
    void skipAttributes()
    {
        const uint64_t tag = take(4), n = take(4);
        if (tag == 0 && n == 0) return;
        if (tag != kAttribute) throw std::runtime_error("NetCDF: bad attribute list in " + mPath);
        for (uint64_t a = 0; a < n; ++a) {
            name();
            const uint32_t t = (uint32_t)take(4);
            const uint64_t cnt = take(4);
            mPos += (std::size_t)((cnt * typeSize(t) + 3) / 4 * 4);
        }
    }
    void parse()
    {
        if (mBytes.size() < 8 || mBytes.compare(0, 3, "CDF") != 0)
            throw std::runtime_error("NetCDF: not a classic NetCDF file: " + mPath);
        mH.version = mBytes[3];
        if (mH.version != 1 && mH.version != 2)
            throw std::runtime_error("NetCDF: only the classic formats CDF-1 / CDF-2 are read: " + mPath);
        mPos = 4;
        mH.numrecs = take(4);
        uint64_t tag = take(4), n = take(4);
        if (!(tag == 0 && n == 0)) {
            if (tag != kDimension) throw std::runtime_error("NetCDF: bad dimension list in " + mPath);
            for (uint64_t d = 0; d < n; ++d) {
                Header::Dim dm;
                dm.name = name();
                dm.len = take(4);
                mH.dims.push_back(dm);
            }
        }
        skipAttributes();  // global attributes
        tag = take(4);
        n = take(4);
        if (!(tag == 0 && n == 0)) {
            if (tag != kVariable) throw std::runtime_error("NetCDF: bad variable list in " + mPath);
            for (uint64_t k = 0; k < n; ++k) {
                Header::Var v;
                v.name = name();
                const uint64_t nd = take(4);
                for (uint64_t d = 0; d < nd; ++d) v.dims.push_back((uint32_t)take(4));
                skipAttributes();
                v.type = (uint32_t)take(4);
                v.vsize = take(4);
                v.begin = take(mH.version == 2 ? 8 : 4);
                mH.vars.push_back(v);
            }
        }
        if (mH.numrecs == 0xFFFFFFFFu) {  // streaming: records up to the end of the file
            const uint64_t rs = mH.recsize();
            uint64_t first = mBytes.size();
            for (const Header::Var& v : mH.vars)
                if (mH.isRecordVar(v) && v.begin < first) first = v.begin;
            mH.numrecs = rs ? (mBytes.size() - first) / rs : 0;
        }
    }
    std::string mPath, mBytes;
    std::size_t mPos = 0;
    Header mH;
};

}  // namespace netcdf

// One snapshot file: <nameDataFile>.nc with the given variables on the
// global grid of nNodes nodes.
template <typename TT, std::size_t DIM>
class DataFile {
public:
    DataFile() : mName("EMPTYFILE.nc") {}
    DataFile(const std::string& nameDataFile, const std::vector<std::string>& nameVariables,
             const Ntuple<int, DIM>& nNodes)
        : mName(nameDataFile + ".nc"), mVars(nameVariables), mNnodes(nNodes)
    {
    }
    const std::string& nameDataFile() const { return mName; }
    const std::vector<std::string>& nameVariables() const { return mVars; }
    int nWritesCurr() const { return mNwrites; }
    std::size_t nNodesTotal() const
    {
        std::size_t n = 1;
        for (std::size_t p = 0; p < DIM; ++p) n *= (std::size_t)mNnodes.elem(p);
        return n;
    }

    // Appends record nWritesCurr(): node block (x fastest) of variable k
    // when writeToFile[k], else the fill value (DataFileApi.hpp:485-585).
    void write(const std::vector<const TT*>& nodeData, const std::vector<bool>& writeToFile)
    {
        if (nodeData.size() != mVars.size() || writeToFile.size() != mVars.size())
            throw std::runtime_error("DataFile::write: one entry per variable expected");
        if (!mIsOpen) create();
        std::fstream f(mName.c_str(), std::ios::in | std::ios::out | std::ios::binary);
        if (!f) throw std::runtime_error("DataFile::write: cannot reopen " + mName);
        const std::size_t nn = nNodesTotal();
        const uint64_t rec = (uint64_t)mNwrites;
        std::vector<double> tmp(nn);
        std::string buf;
        for (std::size_t k = 0; k < mVars.size(); ++k) {
            if (writeToFile[k])
                for (std::size_t i = 0; i < nn; ++i) tmp[i] = static_cast<double>(nodeData[k][i]);
            else
                for (std::size_t i = 0; i < nn; ++i) tmp[i] = netcdf::kFillDouble;
            buf.clear();
            buf.reserve(nn * 8);
            netcdf::putDoublesBE(buf, tmp.data(), nn);
            f.seekp((std::streamoff)(mBegin[k] + rec * mRecSize));
            f.write(buf.data(), (std::streamsize)buf.size());
        }
        ++mNwrites;
        std::string nr;
        netcdf::put32(nr, (uint32_t)mNwrites);
        f.seekp(4);
        f.write(nr.data(), 4);
        if (!f) throw std::runtime_error("DataFile::write: write failed on " + mName);
    }

    // Record 0 of each variable into nodeData[k] (x fastest), after checking
    // the file's grid dimensions (DataFileApi.hpp:588-700).
    void read(const std::vector<TT*>& nodeData) const
    {
        if (nodeData.size() != mVars.size()) throw std::runtime_error("DataFile::read: one entry per variable expected");
        netcdf::Reader r(mName);
        for (std::size_t p = 0; p < DIM; ++p) {
            std::ostringstream oss;
            oss << "nNodes_" << p;
            bool found = false;
            for (const netcdf::Header::Dim& d : r.header().dims)
                if (d.name == oss.str()) {
                    found = true;
                    if ((long)d.len != (long)mNnodes.elem(p))
                        throw std::runtime_error("DataFile::read: grid dimensions of " + mName + " do not fit the grid");
                }
            if (!found) throw std::runtime_error("DataFile::read: " + mName + " has no dimension " + oss.str());
        }
        const std::size_t nn = nNodesTotal();
        for (std::size_t k = 0; k < mVars.size(); ++k) {
            const std::vector<double> v = r.read(mVars[k], 0);
            if (v.size() != nn) throw std::runtime_error("DataFile::read: " + mVars[k] + " has the wrong size");
            for (std::size_t i = 0; i < nn; ++i) nodeData[k][i] = static_cast<TT>(v[i]);
        }
    }

private:
    // Header + "version" (DataFileApi.hpp:392-477): dimensions nNodes_0 ..
    // nNodes_{DIM-1}, time (unlimited); variables version, then the data
    // fields with dimensions (time, nNodes_{DIM-1}, .., nNodes_0).
    void create()
    {
        const uint64_t vsize = (uint64_t)nNodesTotal() * 8;
        mRecSize = vsize * mVars.size();
        int version = 1;
        for (;;) {
            std::string h("CDF");
            h.push_back((char)version);
            netcdf::put32(h, 0);  // numrecs
            netcdf::put32(h, netcdf::kDimension);
            netcdf::put32(h, (uint32_t)(DIM + 1));
            for (std::size_t p = 0; p < DIM; ++p) {
                std::ostringstream oss;
                oss << "nNodes_" << p;
                netcdf::putName(h, oss.str());
                netcdf::put32(h, (uint32_t)mNnodes.elem(p));
            }
            netcdf::putName(h, "time");
            netcdf::put32(h, 0);
            netcdf::put32(h, 0);  // no global attributes
            netcdf::put32(h, 0);
            netcdf::put32(h, netcdf::kVariable);
            netcdf::put32(h, (uint32_t)(1 + mVars.size()));
            // sizes first, offsets patched in once the header length is known
            std::vector<std::size_t> beginAt;
            auto var = [&](const std::string& name, const std::vector<uint32_t>& dims, uint32_t type, uint64_t vs) {
                netcdf::putName(h, name);
                netcdf::put32(h, (uint32_t)dims.size());
                for (uint32_t d : dims) netcdf::put32(h, d);
                netcdf::put32(h, 0);  // no variable attributes
                netcdf::put32(h, 0);
                netcdf::put32(h, type);
                netcdf::put32(h, (uint32_t)std::min<uint64_t>(vs, 0xFFFFFFFFu));
                beginAt.push_back(h.size());
                if (version == 2) netcdf::put64(h, 0);
                else netcdf::put32(h, 0);
            };
            var("version", std::vector<uint32_t>(), netcdf::kInt, 4);
            std::vector<uint32_t> dims(1, (uint32_t)DIM);  // time
            for (std::size_t p = DIM; p-- > 0;) dims.push_back((uint32_t)p);
            for (const std::string& v : mVars) var(v, dims, netcdf::kDouble, vsize);
            const uint64_t hsize = h.size();
            const uint64_t lastBegin = hsize + 4 + (mVars.empty() ? 0 : (mVars.size() - 1) * vsize);
            if (version == 1 && (lastBegin > 0x7FFFFFFFu || vsize > 0xFFFFFFFCu)) {
                version = 2;  // 64-bit offsets
                continue;
            }
            mBegin.assign(mVars.size(), 0);
            std::vector<uint64_t> begins(1, hsize);
            for (std::size_t k = 0; k < mVars.size(); ++k) {
                mBegin[k] = hsize + 4 + k * vsize;
                begins.push_back(mBegin[k]);
            }
            for (std::size_t k = 0; k < begins.size(); ++k) {
                std::string b;
                if (version == 2) netcdf::put64(b, begins[k]);
                else netcdf::put32(b, (uint32_t)begins[k]);
                h.replace(beginAt[k], b.size(), b);
            }
            netcdf::put32(h, (uint32_t)netcdf::kDataFileVersion);  // the version variable's value
            std::ofstream f(mName.c_str(), std::ios::binary | std::ios::trunc);
            if (!f) throw std::runtime_error("DataFile: cannot create " + mName);
            f.write(h.data(), (std::streamsize)h.size());
            if (!f) throw std::runtime_error("DataFile: cannot write " + mName);
            break;
        }
        mIsOpen = true;
    }

    std::string mName;
    std::vector<std::string> mVars;
    Ntuple<int, DIM> mNnodes;
    bool mIsOpen = false;
    int mNwrites = 0;
    uint64_t mRecSize = 0;
    std::vector<uint64_t> mBegin;
};

}  // namespace ScaFES
#endif
