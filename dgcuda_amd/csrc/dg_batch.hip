#include "dg_batch.h"

namespace dgmi {

DGMI_TABLE_UPLOADER(upload_tables_batch)

Status Batch::create(int, Batch**) { return Status::err(DG_ERR_UNSUPPORTED, "batch not built yet"); }

}  // namespace dgmi
