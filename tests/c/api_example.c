/* A program written against the reference's public API (src/virginian.h),
 * built against include/virginian.h and linked with libvirginian_amd.so.
 * Prints the row counts of the reference's nine condition queries
 * (src/test/sql.cc:15-37); with no HIP device it prints "no-device".  The
 * database is backed by the file named on the command line, created afresh
 * as the reference's test fixture does (src/test/test.cc:15-17) and written
 * by virg_close. */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>
#include "virginian.h"
#include "virg_gpu.h"

static const char *queries[9] = {
	"select col0 from test where col0 < 9 and col0 >= 7 or col0 = 3",
	"select col0 from test where col0 >= 8 or col0 = 3 and col0 < 1",
	"select col0 from test where col0 = 8 and col0 = 8",
	"select col0 from test where col0 = 8 or col0 = 8",
	"select col0 from test where col0 = 2",
	"select col0 from test where (col0 = 2)",
	"select col0 from test where col0 < 5 and (col0 = 7 or col0 = 2)",
	"select col0 from test where col0 = 7 or col0 = 2 or col0 = 8",
	"select col0 from test where col0 < 5 and (col0 = 2 or col0 = 8 or col0 = 1)",
};

int main(int argc, char **argv)
{
	const char *dbfile = argc > 1 ? argv[1] : "testdb";
	virginian v;
	if (virg_init(&v) != VIRG_SUCCESS)
		return 1;
	unlink(dbfile);
	if (virg_db_create(&v, dbfile) != VIRG_SUCCESS)
		return 4;
	virg_table_create(&v, "test", VIRG_INT);
	virg_table_addcolumn(&v, 0, "col0", VIRG_INT);
	virg_table_addcolumn(&v, 0, "col1", VIRG_INT);
	virg_table_addcolumn(&v, 0, "col2", VIRG_INT);
	for (int i = 0; i < 10; i++) {
		int row[3] = { i, i + 1, i + 2 };
		virg_table_insert(&v, 0, (char *)&i, (char *)row, NULL);
	}
	unsigned n;
	virg_table_numrows(&v, 0, &n);
	printf("rows %u\n", n);

	/* SQL compiles without a device */
	virg_vm *vm = virg_vm_init();
	if (virg_sql(&v, queries[0], vm) != VIRG_SUCCESS)
		return 2;
	printf("ops %u\n", vm->num_ops);
	virg_vm_cleanup(&v, vm);

	char name[64];
	int cus = 0;
	if (vg_device_info(&v, name, sizeof name, &cus) != VIRG_SUCCESS) {
		printf("no-device\n");
		return virg_close(&v) == VIRG_SUCCESS ? 0 : 5;
	}
	v.use_gpu = 1;
	for (int k = 0; k < 9; k++) {
		virg_reader *r;
		if (virg_query(&v, &r, queries[k]) != VIRG_SUCCESS)
			return 3;
		unsigned rows = 0;
		virg_reader_getrows(&v, r, &rows);
		printf("q%d %u\n", k, rows);
		virg_release(&v, r);
	}
	return virg_close(&v) == VIRG_SUCCESS ? 0 : 5;
}
