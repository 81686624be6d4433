// ScaFES.hpp -- umbrella include of the host API (role of src/ScaFES.hpp:467-474).
//
// A problem class written against ScaFES's Problem/DataField/GridGlobal/
// GridSub/Ntuple/Parameters compiles against these headers unchanged; to
// run, it additionally names its update kernel by specialising
// ScaFES::DeviceTraits (ScaFES_Device.hpp).
#ifndef SCAFES_HOST_SCAFES_HPP_
#define SCAFES_HOST_SCAFES_HPP_

#include "ScaFES_DataField.hpp"
#include "ScaFES_DataFile.hpp"
#include "ScaFES_Device.hpp"
#include "ScaFES_Grid.hpp"
#include "ScaFES_GridGlobal.hpp"
#include "ScaFES_GridSub.hpp"
#include "ScaFES_Ntuple.hpp"
#include "ScaFES_Parameters.hpp"
#include "ScaFES_Problem.hpp"
#include "ScaFES_Timer.hpp"

#endif
