#!/bin/bash
# Writes tests/golden/abi.json: the sizes of the reference's 15 ABI structs
# (the virg_testsizes array of src/virginian.h, as src/test/virgstate.cc
# compares them), its opcode and type numbering and its status codes, all
# read by compiling a probe against the reference's own header.
# Run in the container that holds the reference; the output is committed.
set -euo pipefail
REF=${REF:-/root/reference/src}
# virginian.h includes <cuda.h>, <cuda_runtime_api.h>, <driver_types.h> for
# declarations only; the image carries those headers inside triton.
CUDA_INC=${CUDA_INC:-$(python3 -c 'import triton, os; print(os.path.join(os.path.dirname(triton.__file__), "backends/nvidia/include"))')}
OUT=$(cd "$(dirname "$0")" && pwd)/abi.json
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
cat > "$TMP/probe.c" <<'PROBE'
#include "virginian.h"
#define OPC(x) printf("    \"%s\": %d,\n", #x, (int)x)
int main(void)
{
	unsigned i, n = sizeof(virg_testsizes) / sizeof(size_t);
	printf("{\n  \"struct_sizes\": [");
	for (i = 0; i < n; i++)
		printf("%s%zu", i ? ", " : "", virg_testsizes[i]);
	printf("],\n  \"opcodes\": {\n");
	OPC(OP_Table); OPC(OP_ResultColumn); OPC(OP_Parallel); OPC(OP_Finish);
	OPC(OP_Column); OPC(OP_Rowid); OPC(OP_Result); OPC(OP_Converge);
	OPC(OP_Invalid); OPC(OP_Cast); OPC(OP_Integer); OPC(OP_Float);
	OPC(OP_Le); OPC(OP_Lt); OPC(OP_Ge); OPC(OP_Gt); OPC(OP_Eq); OPC(OP_Neq);
	OPC(OP_Add); OPC(OP_Sub); OPC(OP_Mul); OPC(OP_Div); OPC(OP_And);
	OPC(OP_Or); OPC(OP_Not);
	printf("    \"OP_Nop\": %d\n  },\n", (int)OP_Nop);
	printf("  \"types\": {\"VIRG_INT\": %d, \"VIRG_INT64\": %d, \"VIRG_FLOAT\": %d, "
	       "\"VIRG_DOUBLE\": %d, \"VIRG_CHAR\": %d},\n",
	       VIRG_INT, VIRG_INT64, VIRG_FLOAT, VIRG_DOUBLE, VIRG_CHAR);
	printf("  \"type_sizes\": [%zu, %zu, %zu, %zu, %zu],\n", virg_sizes[0], virg_sizes[1],
	       virg_sizes[2], virg_sizes[3], virg_sizes[4]);
	printf("  \"status\": {\"VIRG_SUCCESS\": %d, \"VIRG_FAIL\": %d},\n", VIRG_SUCCESS, VIRG_FAIL);
	printf("  \"limits\": {\"VIRG_OPS\": %d, \"VIRG_REGS\": %d, \"VIRG_MAX_COLUMNS\": %d, "
	       "\"VIRG_TABLET_SIZE\": %d, \"VIRG_MEM_TABLETS\": %d, \"VIRG_ROW_BUFFER\": %d}\n}\n",
	       VIRG_OPS, VIRG_REGS, VIRG_MAX_COLUMNS, VIRG_TABLET_SIZE, VIRG_MEM_TABLETS,
	       VIRG_ROW_BUFFER);
	return 0;
}
PROBE
gcc -std=gnu99 -w -I "$REF" -I "$CUDA_INC" "$TMP/probe.c" -o "$TMP/probe"
"$TMP/probe" > "$OUT"
echo "wrote $OUT"
