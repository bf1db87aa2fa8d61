#include "dg_batch.h"

namespace dgmi {

Status Batch::create(int, Batch**) { return Status::err(DG_ERR_UNSUPPORTED, "batch not built yet"); }

}  // namespace dgmi
