// Test infrastructure only: the reference's own ItemStream container
// (util/ItemStream.cpp, util/ItemStreamSerialization.cpp, compiled in place by
// oracle/Makefile into oracle/_ref/) driven the way the reference's shape
// library stores a baked TetSolidShape3 (Geo/shape/ShapeLibrary.cpp:332-359:
// one complex item of type eType_ShapeDef holding shape_type, version, the
// four counts and the five arrays as byte blocks). Used to write golden shape
// files and to read files written by the library under test. Never shipped.
#include <util/ItemStream.h>
#include <Geo/ShapeTypes.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace {
const uint16 kTypeShapeDef = uint16(cNumPlaTypes + 1);  // Geo/shape/ShapeLibrary.h:9
const uint32 kTetSolidVersion = 4;                      // TetSolidShape3::cVersion (TetSolidShape3.h:83)
}

extern "C" {

uint32_t ref_is_shape_type_tetsolid3() { return uint32_t(geo::eShape_TetSolid3); }
uint32_t ref_is_type_shapedef() { return kTypeShapeDef; }

// t16: 8 uint16 per tet (4 VID, 4 NTID); bf16: 3 per face; l16: 2 per layer.
int ref_is_write_tetshape(const char* path, int binary, uint32_t nv, const float* pos, uint32_t nt,
                          const uint16_t* t16, uint32_t nbf, const uint16_t* bf16, uint32_t nl,
                          const uint16_t* l16) {
    util::ItemStream is(1 << 10, 1 << 8);
    is.BeginComplex(int32(1), kTypeShapeDef);
    {
        is.Write("shape_type", uint32(geo::eShape_TetSolid3));
        is.Write("version", kTetSolidVersion);
        is.Write("num_v", uint32(nv));
        is.Write("num_t", uint32(nt));
        is.Write("num_bf", uint32(nbf));
        is.Write("num_l", uint32(nl));
        std::vector<uint8> zeros(nv, 0);  // vertex_type: one unused byte each
        is.WriteArray("vec_points", reinterpret_cast<const uint8*>(pos), nv * 12u);
        is.WriteArray("vec_v", zeros.data(), nv);
        is.WriteArray("vec_t", reinterpret_cast<const uint8*>(t16), nt * 16u);
        is.WriteArray("vec_bf", reinterpret_cast<const uint8*>(bf16), nbf * 6u);
        is.WriteArray("vec_l", reinterpret_cast<const uint8*>(l16), nl * 4u);
    }
    is.EndComplex();
    return (binary ? is.SaveBin(path) : is.SaveTxt(path)) ? 1 : 0;
}

// Reads the first complex item holding shape_type == eShape_TetSolid3; the
// caller's buffers must hold cap_* entries. Returns 1 on success.
int ref_is_read_tetshape(const char* path, int binary, uint32_t cap_v, uint32_t cap_t, uint32_t cap_bf,
                         uint32_t cap_l, uint32_t* nv, float* pos, uint32_t* nt, uint16_t* t16, uint32_t* nbf,
                         uint16_t* bf16, uint32_t* nl, uint16_t* l16) {
    util::ItemStream is;
    if (!(binary ? is.LoadBin(path) : is.LoadTxt(path))) return 0;
    for (util::ItemStream::ItemIt it = is.Begin(); it.IsValid(); ++it) {
        if (!it.IsComplex()) continue;
        util::ItemStream::ItemIt sub = it.GetSubItem();
        util::ItemStream::ItemIt st = sub.Find("shape_type");
        if (!st.IsValid() || st.GetNumber<uint32>() != uint32(geo::eShape_TetSolid3)) continue;
        *nv = sub.Find("num_v").GetNumber<uint32>();
        *nt = sub.Find("num_t").GetNumber<uint32>();
        *nbf = sub.Find("num_bf").GetNumber<uint32>();
        *nl = sub.Find("num_l").GetNumber<uint32>();
        if (*nv > cap_v || *nt > cap_t || *nbf > cap_bf || *nl > cap_l) return 0;
        std::memcpy(pos, sub.Find("vec_points").GetArrayPtr<uint8>(), *nv * 12u);
        std::memcpy(t16, sub.Find("vec_t").GetArrayPtr<uint8>(), *nt * 16u);
        std::memcpy(bf16, sub.Find("vec_bf").GetArrayPtr<uint8>(), *nbf * 6u);
        std::memcpy(l16, sub.Find("vec_l").GetArrayPtr<uint8>(), *nl * 4u);
        return 1;
    }
    return 0;
}

}  // extern "C"
