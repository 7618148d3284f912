#pragma once
#include "voreen/core/shim_base.h"
