// Where a paged container holds its arrays (interface of the reference's
// location_enums.h, kept so that callers compile unchanged).
#pragma once

enum class LOCATION { CPU, GPU };

enum class CPU_EXISTENCE { NON_EXISTENT, EXISTENT };

enum class GPU_EXISTENCE { NON_EXISTENT, EXISTENT };
