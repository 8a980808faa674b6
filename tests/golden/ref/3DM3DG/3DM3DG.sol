N_cell: 100
N_dir: 128
N_band: 1
num_theta: 4
num_phi: 4

Iteration #0	 Margin per band per cell: 1.70164e+11
Iteration #1	 Margin per band per cell: 8.86787e+10
Iteration #2	 Margin per band per cell: 3.58285e+10
Iteration #3	 Margin per band per cell: 1.24964e+10
Iteration #4	 Margin per band per cell: 5.76127e+09
Iteration #5	 Margin per band per cell: 3.17729e+09
Iteration #6	 Margin per band per cell: 1.80135e+09
Iteration #7	 Margin per band per cell: 1.09698e+09
Iteration #8	 Margin per band per cell: 6.87615e+08
Iteration #9	 Margin per band per cell: 4.45965e+08
Iteration #10	 Margin per band per cell: 2.96896e+08
Iteration #11	 Margin per band per cell: 2.00858e+08
Iteration #12	 Margin per band per cell: 1.38609e+08
Iteration #13	 Margin per band per cell: 9.69127e+07
Iteration #14	 Margin per band per cell: 6.87031e+07
Iteration #15	 Margin per band per cell: 4.92767e+07
Iteration #16	 Margin per band per cell: 3.57054e+07
Iteration #17	 Margin per band per cell: 2.61156e+07
Iteration #18	 Margin per band per cell: 1.92518e+07
Iteration #19	 Margin per band per cell: 1.42919e+07
Iteration #20	 Margin per band per cell: 1.06734e+07
Iteration #21	 Margin per band per cell: 8.0126e+06
Iteration #22	 Margin per band per cell: 6.04204e+06
Iteration #23	 Margin per band per cell: 4.57362e+06
Iteration #24	 Margin per band per cell: 3.4735e+06
Iteration #25	 Margin per band per cell: 2.64553e+06
Iteration #26	 Margin per band per cell: 2.01983e+06
Iteration #27	 Margin per band per cell: 1.54537e+06
Iteration #28	 Margin per band per cell: 1.1845e+06
Iteration #29	 Margin per band per cell: 909309
Iteration #30	 Margin per band per cell: 698980
Iteration #31	 Margin per band per cell: 537921
Iteration #32	 Margin per band per cell: 414377
Iteration #33	 Margin per band per cell: 319477
Iteration #34	 Margin per band per cell: 246486
Iteration #35	 Margin per band per cell: 190290
Iteration #36	 Margin per band per cell: 146982
Iteration #37	 Margin per band per cell: 113582
Iteration #38	 Margin per band per cell: 87804.3
Iteration #39	 Margin per band per cell: 67899.4
Iteration #40	 Margin per band per cell: 52520.8
Iteration #41	 Margin per band per cell: 40635
Iteration #42	 Margin per band per cell: 31444.9
Iteration #43	 Margin per band per cell: 24337.5
Iteration #44	 Margin per band per cell: 18839
Iteration #45	 Margin per band per cell: 14584.5
Iteration #46	 Margin per band per cell: 11291.8
Iteration #47	 Margin per band per cell: 8743.3
Iteration #48	 Margin per band per cell: 6770.37
Iteration #49	 Margin per band per cell: 5242.94
Iteration #50	 Margin per band per cell: 4060.27
Iteration #51	 Margin per band per cell: 3144.51
Iteration #52	 Margin per band per cell: 2435.35
Iteration #53	 Margin per band per cell: 1886.17
Iteration #54	 Margin per band per cell: 1460.86
Iteration #55	 Margin per band per cell: 1131.47
Iteration #56	 Margin per band per cell: 876.355
Iteration #57	 Margin per band per cell: 678.77
Iteration #58	 Margin per band per cell: 525.734
Iteration #59	 Margin per band per cell: 407.205
Iteration #60	 Margin per band per cell: 315.398
Iteration #61	 Margin per band per cell: 244.291
Iteration #62	 Margin per band per cell: 189.214
Iteration #63	 Margin per band per cell: 146.555
Iteration #64	 Margin per band per cell: 113.513
Iteration #65	 Margin per band per cell: 87.9211
Iteration #66	 Margin per band per cell: 68.0983
Iteration #67	 Margin per band per cell: 52.745
Iteration #68	 Margin per band per cell: 40.8528
Iteration #69	 Margin per band per cell: 31.642
Iteration #70	 Margin per band per cell: 24.5077
Iteration #71	 Margin per band per cell: 18.9821
Iteration #72	 Margin per band per cell: 14.7021
Iteration #73	 Margin per band per cell: 11.3872
Iteration #74	 Margin per band per cell: 8.81967
Iteration #75	 Margin per band per cell: 6.83105
Iteration #76	 Margin per band per cell: 5.29079
Iteration #77	 Margin per band per cell: 4.09783
Iteration #78	 Margin per band per cell: 3.17384
Iteration #79	 Margin per band per cell: 2.4582
Iteration #80	 Margin per band per cell: 1.90391
Iteration #81	 Margin per band per cell: 1.47461
Iteration #82	 Margin per band per cell: 1.1421
Iteration #83	 Margin per band per cell: 0.884574
Iteration #84	 Margin per band per cell: 0.685112
Iteration #85	 Margin per band per cell: 0.530628
Iteration #86	 Margin per band per cell: 0.410976
Iteration #87	 Margin per band per cell: 0.318306
Iteration #88	 Margin per band per cell: 0.24653
Iteration #89	 Margin per band per cell: 0.19094
Iteration #90	 Margin per band per cell: 0.147884
Iteration #91	 Margin per band per cell: 0.114538
Iteration #92	 Margin per band per cell: 0.0887104
Iteration #93	 Margin per band per cell: 0.0687071
Iteration #94	 Margin per band per cell: 0.053214
Iteration #95	 Margin per band per cell: 0.0412148
Iteration #96	 Margin per band per cell: 0.0319211
Iteration #97	 Margin per band per cell: 0.0247232
Iteration #98	 Margin per band per cell: 0.0191483
Iteration #99	 Margin per band per cell: 0.0148305
Iteration #100	 Margin per band per cell: 0.0114864
Iteration #101	 Margin per band per cell: 0.00889631
Iteration #102	 Margin per band per cell: 0.00689026
Iteration #103	 Margin per band per cell: 0.00533658
Iteration #104	 Margin per band per cell: 0.00413323
Iteration #105	 Margin per band per cell: 0.00320123
Iteration #106	 Margin per band per cell: 0.00247939
Iteration #107	 Margin per band per cell: 0.00192032
Iteration #108	 Margin per band per cell: 0.0014873
Iteration #109	 Margin per band per cell: 0.00115194
Iteration #110	 Margin per band per cell: 0.000892188
Iteration #111	 Margin per band per cell: 0.000691013
Iteration #112	 Margin per band per cell: 0.000535198
Iteration #113	 Margin per band per cell: 0.00041452
Iteration #114	 Margin per band per cell: 0.000321051
Iteration #115	 Margin per band per cell: 0.000248659
Iteration #116	 Margin per band per cell: 0.00019259
Iteration #117	 Margin per band per cell: 0.000149165
Iteration #118	 Margin per band per cell: 0.00011553
Iteration #119	 Margin per band per cell: 8.94806e-05

heat flux of Boundary Condition #0: -9.87072e+08
heat flux of Boundary Condition #1: -125980
heat flux of Boundary Condition #2: -586552
heat flux of Boundary Condition #3: 340058
heat flux of Boundary Condition #4: 303519
heat flux of Boundary Condition #5: 9.86886e+08
