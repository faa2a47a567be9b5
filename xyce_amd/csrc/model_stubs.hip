// Placeholders for model back ends not yet built: group creation fails
// with XA_ERR_UNSUPPORTED.
#include "xa_internal.h"
namespace xa {
const ModelImpl* model_bjt() { return nullptr; }
}  // namespace xa
