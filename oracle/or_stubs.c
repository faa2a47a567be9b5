/* Placeholders for oracle models not yet restated (test infrastructure). */
#include "or_internal.h"
const or_model_ops or_bjt_ops = {"bjt"};
