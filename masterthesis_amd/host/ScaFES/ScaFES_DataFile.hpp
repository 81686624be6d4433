// ScaFES_DataFile.hpp -- output policy of a data field.
//
// The enum values are the reference's (src/ScaFES_DataFile.hpp:24-32).
// NetCDF snapshot output is not part of this engine (no libnetcdf on the
// target image); the policy is accepted and recorded.
#ifndef SCAFES_HOST_DATAFILE_HPP_
#define SCAFES_HOST_DATAFILE_HPP_

namespace ScaFES {

enum class WriteHowOften {
    NEVER = 1,
    AT_START = 2,
    AT_END = 3,
    AT_START_AND_END = 4,
    LIKE_GIVEN_AT_CL = 8,
    ALWAYS = 16,
};

}  // namespace ScaFES
#endif
