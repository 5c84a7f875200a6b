// sv_program.hip -- placeholder translation unit for the gate-program scheduler.
#include "sv_program.h"
